// FP32-bulk RK stage with p-advection and WENO5 in p, one workgroup per
// q-cell, with the p-fluxes computed by sweeps along p-lines.
//
// qcell_kernel evaluates the three p-flux differences of every phase-space
// cell on its own: per axis, two face states from four WENO indicators,
// gathered from a seven-cell window with periodic index arithmetic.  Here
// each p-axis is swept instead: a thread marches along a segment of one
// p-line of the q-cell's block (in LDS), computing each face state once and
// each cell's indicator once, and the flux difference of every cell is added
// into an LDS accumulator.  The q-fluxes, BGK source, RK combination and
// moments then run over the p-cells in natural order (a wave's loads and
// stores are contiguous), with the upwind side of each q-face selected per
// lane (qs_q_flux_diff), taking the p-part from the accumulator.
//
// C_p along its own axis does not depend on that axis's velocity
// (advection_terms.py:147-188), so C is constant along a p-line and one sign
// per line selects the upwind state: the right face of cell j (window
// j-2..j+2) when C > 0, else the left face of cell j+1 (window j-1..j+3).
// The reference's left-face formula on a window z0..z4 equals its right-face
// formula on the reversed window z4..z0 with the three indicators reversed
// (weno5.py:73-96; the indicators themselves keep their forward formula,
// weno5.py:58-71), so a line with C <= 0 is marched backwards with the
// right-face formula, and lines of either sign share a wave without
// divergence.  LLF equals upwind for this linear flux (C is the same on both
// sides of every face).
//
// LDS: the q-cell's block with each p1-line padded to n_p1 + 1 floats (an odd
// pitch for n_p1 = 24, so a wave's 64 lines hit 64 distinct banks), the
// accumulator in the same padded layout, and the Maxwellian factor tables.
//
// Reference: df_dt_fvm.py:251-323 (p fluxes), weno5.py:57-98,
// riemann.py:46-49; the rest as qcell_kernel.
#pragma once
#include "qcell.h"

namespace bx {

constexpr int QS_SEG = 8;  // cells per sweep segment

// x / d and x % d through the FP32 reciprocal (fvm_math.h).
__device__ __forceinline__ int qs_divmod(int x, int d, float rcp_d, int& rem) {
  return fdivmod(x, d, rcp_d, rem);
}

// The reference's WENO5 indicator (sum-of-squares form of weno5.py:58-71) as
// kappa = 1 / (16 beta^2), so a weight d/16/beta^2 is d * kappa.
__device__ __forceinline__ float qs_kappa(float a, float b, float c) {
  const float s1 = __builtin_fmaf(-2.0f, b, a) + c;
  const float s2 = __builtin_fmaf(3.0f, c, __builtin_fmaf(-4.0f, b, a));
  const float e = 1e-17f * (((1.0f + aabs(a)) + aabs(b)) + aabs(c));
  const float bt = __builtin_fmaf((13.0f / 12.0f) * s1, s1, __builtin_fmaf(0.25f * s2, s2, e));
  return __builtin_amdgcn_rcpf(16.0f * bt * bt);
}

// Right-face WENO5 state of the centre of y0..y4 with indicator quantities
// k1, k2, k3 of the three substencils (weights 1, 10, 5 in units of 1/16).
__device__ __forceinline__ float qs_face(float y0, float y1, float y2, float y3, float y4,
                                         float k1, float k2, float k3) {
  const float u1 = __builtin_fmaf(1.875f, y2, __builtin_fmaf(-1.25f, y1, 0.375f * y0));
  const float u2 = __builtin_fmaf(0.375f, y3, __builtin_fmaf(0.75f, y2, -0.125f * y1));
  const float u3 = __builtin_fmaf(-0.125f, y4, __builtin_fmaf(0.75f, y3, 0.375f * y2));
  const float w2 = 10.0f * k2, w3 = 5.0f * k3;
  const float num = __builtin_fmaf(w3, u3, __builtin_fmaf(w2, u2, k1 * u1));
  const float den = (k1 + w2) + w3;
  return num * __builtin_amdgcn_rcpf(den);
}

// Flux difference C*(s(i+1/2) - s(i-1/2)) along a q axis for the cell at the
// centre of a[k] = f[i-3+k], with the upwind side chosen per lane: both faces
// are right-face reconstructions of selected windows, the lower face of
// window a[0..4] (C > 0: right face of cell i-1) or of the reversed a[5..1]
// (C <= 0: left face of cell i), the upper face of a[1..5] or of the reversed
// a[6..2].  PPM, minmod and piecewise-constant faces are mirror-symmetric
// (ppm.py:83-98: of the two corrections at most one applies); for WENO5 the
// forward indicators are assigned in reverse order (weno5.py:58-96).  For
// the other reconstructions both windows are cut from one upwind-ordered
// window W[k] = a[k] (C > 0) or a[6-k]: the faces of W0..W4 and W1..W5 are
// the lower and upper faces (C > 0) or the upper and lower ones, and PPM's
// limited slopes of W1..W4 serve both.  The lower face state is returned in
// ``lower`` (f_left / f_bot of df_dt_fvm.py:79-80).
template <int RQ>
__device__ __forceinline__ float qs_q_flux_diff(const float a[7], float C, bool pos, float& lower) {
  float sL, sU;
  if constexpr (RQ == RC_WENO5) {
    float L[5], U[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      L[k] = pos ? a[k] : a[5 - k];
      U[k] = pos ? a[k + 1] : a[6 - k];
    }
    float b[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) b[k] = weno_beta<float>(a[k], a[k + 1], a[k + 2]);
    sL = weno_combine(L[0], L[1], L[2], L[3], L[4], pos ? b[0] : b[3], pos ? b[1] : b[2],
                      pos ? b[2] : b[1], true);
    sU = weno_combine(U[0], U[1], U[2], U[3], U[4], pos ? b[1] : b[4], pos ? b[2] : b[3],
                      pos ? b[3] : b[2], true);
  } else {
    float W[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) W[k] = pos ? a[k] : a[6 - k];
    float F0, F1;
    if constexpr (RQ == RC_PPM) {
      float dq[4], l;
#pragma unroll
      for (int k = 0; k < 4; ++k) dq[k] = ppm_dq(W[k], W[k + 1], W[k + 2]);
      ppm_faces_dq(W[1], W[2], W[3], dq[0], dq[1], dq[2], l, F0);
      ppm_faces_dq(W[2], W[3], W[4], dq[1], dq[2], dq[3], l, F1);
    } else {
      F0 = face_value<RQ>(W[0], W[1], W[2], W[3], W[4], true);
      F1 = face_value<RQ>(W[1], W[2], W[3], W[4], W[5], true);
    }
    sL = pos ? F0 : F1;
    sU = pos ? F1 : F0;
  }
  lower = sL;
  return C * sU - C * sL;
}

// Sweep of one p-axis: every line (the other two p-indices fixed) is cut into
// segments of QS_SEG cells; thread t takes task t, t+NT, ... with consecutive
// tasks on consecutive lines of one segment index.  acc[cell] = (or +=)
// (F(j+1/2) - F(j-1/2)) * inv_dp.
template <int AX, int NT>
__device__ inline void qs_sweep(const StageParams& sp, const QCell<float>& c, const float* blk,
                                float* acc, int pitch1) {
  const int n = AX == 0 ? sp.n_p1 : AX == 1 ? sp.n_p2 : sp.n_p3;
  if (n <= 1) return;
  const int stride = AX == 0 ? 1 : AX == 1 ? pitch1 : pitch1 * sp.n_p2;
  // the two other axes: (ia, na) fastest across lines, then (ib, nb)
  const int na = AX == 0 ? sp.n_p2 : sp.n_p1;
  const int nb = AX == 2 ? sp.n_p2 : sp.n_p3;
  const int nlines = na * nb;
  const int nseg = (n + QS_SEG - 1) / QS_SEG;
  const float rcp_lines = 1.0f / float(nlines), rcp_na = 1.0f / float(na);
  const double em = sp.charge / sp.mass;
  const float inv_dp = float(AX == 0 ? sp.inv_dp1 : AX == 1 ? sp.inv_dp2 : sp.inv_dp3);
  for (int t = threadIdx.x; t < nlines * nseg; t += NT) {
    int line, ia;
    const int seg = qs_divmod(t, nlines, rcp_lines, line);
    const int ib = qs_divmod(line, na, rcp_na, ia);
    // velocities of the fixed indices and the line's C (cell-centred fields)
    int base;
    double C;
    if (AX == 0) {
      const double v2 = centre(sp.p2s, ia, sp.dp2), v3 = centre(sp.p3s, ib, sp.dp3);
      C = em * (c.E[0] + v2 * c.B[0][2] - v3 * c.B[0][1]);
      base = pitch1 * (ia + sp.n_p2 * ib);
    } else if (AX == 1) {
      const double v1 = centre(sp.p1s, ia, sp.dp1), v3 = centre(sp.p3s, ib, sp.dp3);
      C = em * (c.E[1] + v3 * c.B[1][0] - v1 * c.B[1][2]);
      base = ia + pitch1 * sp.n_p2 * ib;
    } else {
      const double v1 = centre(sp.p1s, ia, sp.dp1), v2 = centre(sp.p2s, ib, sp.dp2);
      C = em * (c.E[2] + v1 * c.B[2][1] - v2 * c.B[2][0]);
      base = ia + pitch1 * ib;
    }
    const float Cf = float(C);
    const bool pos = Cf > 0.0f;
    // The line is marched in its upwind direction: march cell m is line cell
    // m (C > 0) or n-1-m, so the upwind state of the face above march cell m
    // is always the right-face formula on march cells m-2 .. m+2 with the
    // indicators of m-1, m, m+1 (for C <= 0 that is the left face of line
    // cell n-1-m on its reversed window).
    const int j0 = seg * QS_SEG;
    const int len = n - j0 < QS_SEG ? n - j0 : QS_SEG;
    const int dstr = pos ? stride : -stride;
    // window y[k] = march cell j0-4+k, k = 0 .. QS_SEG+6, periodic along the
    // line (a short last segment loads cells it does not use); the line offset
    // moves by one stride per cell and wraps at the line's ends
    float y[QS_SEG + 7];
    {
      int j = pos ? j0 - 4 : n + 3 - j0;
      j = n >= 4 ? (j < 0 ? j + n : (j >= n ? j - n : j)) : ((j % n) + n) % n;
      int off = j * stride;
      const int end = n * stride;
      const int wrap_at = pos ? end : -stride, wrap_to = pos ? 0 : end - stride;
#pragma unroll
      for (int k = 0; k < QS_SEG + 7; ++k) {
        y[k] = blk[base + off];
        off += dstr;
        off = off == wrap_at ? wrap_to : off;
      }
    }
    // indicator quantity kp[k] of march cell j0-3+k, k = 0 .. len+4: the
    // forward formula on the three cells in line order
    float kp[QS_SEG + 5];
#pragma unroll
    for (int k = 0; k < QS_SEG + 5; ++k) {
      const float lo = pos ? y[k] : y[k + 2], hi = pos ? y[k + 2] : y[k];
      kp[k] = qs_kappa(lo, y[k + 1], hi);
    }
    // flux through the face below march cell j0 (window y[1..5], indicators
    // kp[1..3])
    float Flo = Cf * qs_face(y[1], y[2], y[3], y[4], y[5], kp[1], kp[2], kp[3]);
    int at = base + (pos ? j0 : n - 1 - j0) * stride;
#pragma unroll
    for (int s = 0; s < QS_SEG; ++s) {
      if (s < len) {
        // face above march cell j0+s: window y[s+2 .. s+6], indicators
        // kp[s+2 .. s+4]; the line cell's flux difference is the march
        // difference, negated when the march runs backwards
        const float Fhi = Cf * qs_face(y[s + 2], y[s + 3], y[s + 4], y[s + 5], y[s + 6],
                                       kp[s + 2], kp[s + 3], kp[s + 4]);
        const float dflux = (pos ? Fhi - Flo : Flo - Fhi) * inv_dp;
        if (AX == 0) acc[at] = dflux;
        else acc[at] = acc[at] + dflux;
        Flo = Fhi;
        at += dstr;
      }
    }
  }
}

template <int RQ, int NT>
__global__ __launch_bounds__(NT) void qsweep_kernel(QCParams qp) {
  constexpr int NW = NT / 64;
  extern __shared__ unsigned char smem[];
  __shared__ double red[8][NW];
  const StageParams& sp = qp.sp;
  const int pitch1 = sp.n_p1 + 1;
  const int PP = pitch1 * sp.n_p2 * sp.n_p3;
  float* tab = reinterpret_cast<float*>(smem);
  float* blk = tab + (sp.n_p1 + sp.n_p2 + sp.n_p3);
  float* acc = blk + PP;
  const int q = qc_order(sp, blockIdx.x, gridDim.x);
  QCell<float> c;
  qcell_neighbours(sp, q, c);
  c.pref = 0.0;
  c.e1 = c.e2 = c.e3 = tab;
  if (sp.source) {
    const BgkCell b = qcell_bgk(sp, q);
    c.pref = maxwell_prefactor(b, sp.mass, sp.kB, sp.p_dim);
    float* e1 = tab;
    float* e2 = tab + sp.n_p1;
    float* e3 = e2 + sp.n_p2;
    for (int i = threadIdx.x; i < sp.n_p1; i += blockDim.x)
      e1[i] = float(maxwell_factor(centre(sp.p1s, i, sp.dp1), b.u1, b.T, sp.mass, sp.kB));
    for (int i = threadIdx.x; i < sp.n_p2; i += blockDim.x)
      e2[i] = float(maxwell_factor(centre(sp.p2s, i, sp.dp2), b.u2, b.T, sp.mass, sp.kB));
    for (int i = threadIdx.x; i < sp.n_p3; i += blockDim.x)
      e3[i] = float(maxwell_factor(centre(sp.p3s, i, sp.dp3), b.u3, b.T, sp.mass, sp.kB));
    c.e1 = e1;
    c.e2 = e2;
    c.e3 = e3;
  }
  const float* u = (const float*)sp.u;
  const long P = sp.P;
  const long base = (long)q * P;
  qcell_fields(sp, c);
  // the block into the padded layout (p1-lines of pitch n_p1 + 1); the
  // accumulator is written by the p1 sweep, so it needs no clearing
  for (int p = threadIdx.x; p < sp.P; p += NT) {
    const int j1 = p % sp.n_p1, rest = p / sp.n_p1;
    blk[j1 + pitch1 * rest] = u[base + p];
  }
  if (sp.n_p1 <= 1)
    for (int i = threadIdx.x; i < PP; i += NT) acc[i] = 0.0f;
  __syncthreads();
  qs_sweep<0, NT>(sp, c, blk, acc, pitch1);
  __syncthreads();
  qs_sweep<1, NT>(sp, c, blk, acc, pitch1);
  __syncthreads();
  qs_sweep<2, NT>(sp, c, blk, acc, pitch1);
  __syncthreads();

  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const float* wv = (const float*)sp.w;
  const float* unv = (const float*)sp.un;
  float* outv = (float*)sp.out;
  const bool use_un = sp.a != 0.0;
  const bool w_is_u = sp.w == sp.u;
  const float ta = float(sp.a), tb = float(sp.b), tcdt = float(sp.cdt);
  const float tinv1 = float(sp.inv_dq1), tinv2 = float(sp.inv_dq2), tinvtau = float(sp.inv_tau);
  constexpr int r = recon_reach(RQ) + 1;
  long nb1[7], nb2[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    nb1[k] = (long)c.nb1[k] * P;
    nb2[k] = (long)c.nb2[k] * P;
  }
  double mac[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  double jac[3] = {0.0, 0.0, 0.0};
  const float rcp1 = 1.0f / float(sp.n_p1), rcp2 = 1.0f / float(sp.n_p2);
  // p = threadIdx.x + k NT: its indices (j1, j2, j3) advance by the uniform
  // steps of NT = q1 n1 + d1, q1 = q2 n2 + d2, with one carry each
  const int n1 = sp.n_p1, n2 = sp.n_p2;
  const int q1 = NT / n1, d1 = NT - q1 * n1;
  const int q2 = q1 / n2, d2 = q1 - q2 * n2;
  int j1, j2, j3;
  j3 = qs_divmod(qs_divmod((int)threadIdx.x, n1, rcp1, j1), n2, rcp2, j2);
  // natural p order: a wave's loads and stores are contiguous runs of 64
  for (int p = threadIdx.x; p < sp.P; p += NT) {
    const int j23 = j2 + n2 * j3;
    const double v1 = centre(sp.p1s, j1, sp.dp1);
    const double v2 = centre(sp.p2s, j2, sp.dp2);
    const double v3 = centre(sp.p3s, j3, sp.dp3);
    const int at = j1 + pitch1 * j23;
    const float f = blk[at];
    float t1 = 0.0f, t2 = 0.0f, s1 = f, s2 = f;
    if (sp.n_q1 > 1) {
      float a[7];
#pragma unroll
      for (int k = 0; k < 7; ++k) a[k] = f;
#pragma unroll
      for (int k = 3 - r; k <= 3 + r; ++k)
        if (k != 3) a[k] = u[nb1[k] + p];
      t1 = qs_q_flux_diff<RQ>(a, float(v1), float(v1) > 0.0f, s1);
    }
    if (sp.n_q2 > 1) {
      float a[7];
#pragma unroll
      for (int k = 0; k < 7; ++k) a[k] = f;
#pragma unroll
      for (int k = 3 - r; k <= 3 + r; ++k)
        if (k != 3) a[k] = u[nb2[k] + p];
      t2 = qs_q_flux_diff<RQ>(a, float(v2), float(v2) > 0.0f, s2);
    }
    if (sp.Jout) {
      jac[0] += double(s1) * v1;
      jac[1] += double(s2) * v2;
      jac[2] += double(f) * v3;
    }
    float d = -t1 * tinv1 - t2 * tinv2;
    if (sp.source) {
      float fmb = float(c.pref) * c.e1[j1];
      if (sp.p_dim >= 2) fmb = fmb * c.e2[j2];
      if (sp.p_dim >= 3) fmb = fmb * c.e3[j3];
      d = d + (fmb - f) * tinvtau;
    }
    d = d - acc[at];
    const float wval = w_is_u ? f : wv[base + p];
    const float tv = wval + d * tcdt;
    const float x = use_un ? ta * unv[base + p] + tb * tv : (sp.b == 1.0 ? tv : tb * tv);
    outv[base + p] = x;
    const double xd = double(x);
    mac[0] += xd;
    mac[1] += xd * v1;
    mac[2] += xd * v2;
    mac[3] += xd * v3;
    mac[4] += (0.5 * xd) * ((v1 * v1 + v2 * v2) + v3 * v3);
    j1 += d1;
    const int c1 = j1 >= n1 ? 1 : 0;
    j1 -= c1 ? n1 : 0;
    j2 += d2 + c1;
    const int c2 = j2 >= n2 ? 1 : 0;
    j2 -= c2 ? n2 : 0;
    j3 += q2 + c2;
  }
  qc_finish<NW>(sp, q, mac, jac, red, wave, lane);
}

}  // namespace bx
