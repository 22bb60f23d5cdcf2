// The nonlinear FVM stage in the q1-line layout [q2][p][q1] (q1 fastest), for
// 2-D or 1-D q-space with fields off -- the layout the solver picks when the
// q1 extent is a multiple of 64.
//
// qline_kernel: a wave owns one p-cell and the whole periodic q1 line (NR
// cells per lane, lane l holding q1 = l + 64 k), and marches along q2 over a
// segment of rows:
//  * q2: a 7-row register window per cell; the WENO indicator and the upwind
//    state of each q2 face are computed once and carried to the next row;
//  * q1: the row, its indicators and its face states go through a
//    wave-private LDS line (no workgroup barrier), so every q1 indicator and
//    face is computed once;
//  * one p per wave makes the upwind direction of both axes a scalar branch,
//    and every global access is a coalesced run along q1;
//  * BGK reads the q-cell's Maxwellian from tables [j][q] (prefactor folded
//    into the p1 table), built per stage from the moment block of the input.
// The moments of the stage output are a separate streaming reduction
// (mline_kernel + ml_finalize_kernel): the stencil kernel does no cross-wave
// reduction at all.
//
// Reference: df_dt_fvm.py:56-97, riemann.py:46-49, weno5.py:57-98,
// collision_operator.py:52-65 and 102-116, moments.py:5-22.
#pragma once
#include "qcell.h"

namespace bx {

struct QLParams {
  int n1, n2, P, nseg, n_p1, n_p2, n_p3;  // nseg q2 segments of near-equal length
  const void* u;
  const void* w;
  const void* un;
  void* out;
  double a, b, cdt, inv_dq1, inv_dq2, inv_tau;
  int source, p_dim;
  double p1s, dp1, p2s, dp2, p3s, dp3;
  const void* tab;  // [n_p1 + n_p2 + n_p3][Q] of T: prefactor*e1, e2, e3
};

// WENO5 indicator quantity kept per cell (computed once per cell and axis,
// shared by the three faces that use it): FP64 keeps beta^2 of the
// reference's expanded indicator (weights are formed as products of the
// other two, one division per face); FP32 keeps 1/(16 beta^2) of the
// sum-of-squares indicator (weno5.py:58-80), whose squares would underflow in
// products.  Other reconstructions keep nothing.
template <int R, typename T>
__host__ __device__ inline T ql_kappa(T a, T b, T c) {
  if constexpr (R != RC_WENO5) {
    return T(0);
  } else if constexpr (sizeof(T) == 8) {
    const double bt = weno_beta_literal(a, b, c);
    return bt * bt;
  } else {
    const float bt = weno_beta_sos(a, b, c);
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(16.0f * bt * bt);
#else
    return 1.0f / (16.0f * bt * bt);
#endif
  }
}

// Face state of the centre cell y2 of the window y0..y4: its right face
// (RIGHT) or its left face; ka, kb, kc are the indicator quantities of the
// cells y1, y2, y3.  WENO5 (weno5.py:73-96): the reference's weights
// w = d / (16 beta^2) with d = (1, 10, 5) (right) or (5, 10, 1) (left) in
// units of 1/16; the d factors of the numerator are folded into the
// substencil coefficients (all exact in binary).
template <int R, bool RIGHT, typename T>
__host__ __device__ inline T ql_face(T y0, T y1, T y2, T y3, T y4, T ka, T kb, T kc) {
  if constexpr (R == RC_WENO5) {
    T u1, u2, u3;
    if constexpr (RIGHT) {
      u1 = T(0.375) * y0 - T(1.25) * y1 + T(1.875) * y2;
      u2 = T(-1.25) * y1 + T(7.5) * y2 + T(3.75) * y3;
      u3 = T(1.875) * y2 + T(3.75) * y3 - T(0.625) * y4;
    } else {
      u1 = T(-0.625) * y0 + T(3.75) * y1 + T(1.875) * y2;
      u2 = T(3.75) * y1 + T(7.5) * y2 - T(1.25) * y3;
      u3 = T(1.875) * y2 - T(1.25) * y3 + T(0.375) * y4;
    }
    T w1, w2, w3;
    if constexpr (sizeof(T) == 8) {
      w1 = kb * kc;
      w2 = ka * kc;
      w3 = ka * kb;
    } else {
      w1 = ka;
      w2 = kb;
      w3 = kc;
    }
    const T num = w1 * u1 + w2 * u2 + w3 * u3;
    const T den = RIGHT ? (w1 + T(10) * w2) + T(5) * w3 : (T(5) * w1 + T(10) * w2) + w3;
    return num * fast_rcp(den);
  } else {
    return face_value<R>(y0, y1, y2, y3, y4, RIGHT);
  }
}

// Halo of the LDS q1 lines: cells -3..-1 and n1..n1+2 hold periodic copies,
// so every neighbour read is the lane's own address plus a constant.
constexpr int QL_HALO = 3;

// Store v for cell i of a halo-padded line (n1 >= QL_HALO), with its
// periodic copies.
template <typename T>
__device__ inline void ql_put(T* line, int i, bool live, int n1, T v) {
  if (!live) return;
  line[i] = v;
  if (i < QL_HALO) line[i + n1] = v;
  if (i >= n1 - QL_HALO) line[i - n1] = v;
}

// Stage forms the march is specialised for (out = a*un + b*(w + cdt*L(u))):
// QL_WU_PLAIN  w == u, a == 0, b == 1  (first stage of either stepper)
// QL_WU_COMB   w == u, a != 0          (SSP-RK3 stages 2 and 3)
// QL_W_PLAIN   w != u, a == 0, b == 1  (second stage of the RK2 midpoint)
// QL_GENERAL   anything else, decided at run time.
enum QlMode : int { QL_WU_PLAIN = 0, QL_WU_COMB = 1, QL_W_PLAIN = 2, QL_GENERAL = 3 };

// Every f-sized stream (window rows of u, w, un, the output) is read or
// written once per launch, so it goes through non-temporal accesses and the
// L2 keeps the Maxwellian table rows the p-cells of a segment share.
//
// One wave's march over a q2 segment for the upwind directions P1 (C_q1 > 0)
// and P2 (C_q2 > 0), fixed at compile time, in stage form MODE with PD
// Maxwellian factors per cell (0 = no source; -1 = decided at run time).  Along q2 the register window
// holds five rows y0..y4: rows r-2..r+2 (P2: the face r+1/2 is the right face
// of row r) or r-1..r+3 (not P2: it is the left face of row r+1); either way
// the face uses the window's centre y2 and the indicators of y1, y2, y3, and
// the state of face r-1/2 is carried from the previous row.  Along q1 the row
// goes through the wave's LDS line: values, then indicators, then the state
// at each cell's lower face.
template <typename T, int R, int NR, bool P1, bool P2, int MODE, int PD>
__device__ __forceinline__ void ql_march(const QLParams& q, int p, int seg, int lane, T* fl, T* kl,
                                         T* sl) {
  const int n1 = q.n1, n2 = q.n2;
  const int j1 = p % q.n_p1, j23 = p / q.n_p1, j2 = j23 % q.n_p2, j3 = j23 / q.n_p2;
  const T C1 = T(centre(q.p1s, j1, q.dp1));
  const T C2 = T(centre(q.p2s, j2, q.dp2));
  const long rowP = (long)q.P * n1;  // elements per q2 row
  // segment seg covers rows [seg*n2/nseg, (seg+1)*n2/nseg)
  const int r0 = (int)((long)seg * n2 / q.nseg);
  const int len = (int)((long)(seg + 1) * n2 / q.nseg) - r0;
  const T* u = (const T*)q.u + (long)p * n1;
  const T* wv = (const T*)q.w + (long)p * n1;
  const T* unv = (const T*)q.un + (long)p * n1;
  T* outv = (T*)q.out + (long)p * n1;
  const long Q = (long)n1 * n2;
  const T* t1p = (const T*)q.tab + (long)j1 * Q;
  const T* t2p = (const T*)q.tab + (long)(q.n_p1 + j2) * Q;
  const T* t3p = (const T*)q.tab + (long)(q.n_p1 + q.n_p2 + j3) * Q;
  const bool w_is_u = MODE == QL_WU_PLAIN || MODE == QL_WU_COMB ? true
                      : MODE == QL_W_PLAIN ? false : q.w == q.u;
  const bool use_un = MODE == QL_WU_COMB ? true : MODE == QL_GENERAL ? q.a != 0.0 : false;
  const bool plain = MODE == QL_WU_PLAIN || MODE == QL_W_PLAIN ? true
                     : MODE == QL_WU_COMB ? false : q.b == 1.0;
  const T ta = T(q.a), tb = T(q.b), tcdt = T(q.cdt);
  const T tinv1 = T(q.inv_dq1), tinv2 = T(q.inv_dq2), tinvtau = T(q.inv_tau);
  const bool do_q1 = n1 > 1, do_q2 = n2 > 1;
  const int pd = PD >= 0 ? PD : (q.source ? q.p_dim : 0);
  const bool src = pd >= 1, src2 = pd >= 2, src3 = pd >= 3;
  constexpr int B = P2 ? -2 : -1;  // window row offset from the current row

  // lane cells; their periodic q1 neighbours are read from the halo-padded
  // LDS lines at constant offsets
  int i1[NR];
  bool live[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    i1[k] = lane + 64 * k;
    live[k] = i1[k] < n1;
    if (!live[k]) i1[k] = 0;
  }

  T y[NR][5], ka[NR], kb[NR], s2m[NR], wc[NR], unc[NR], tc1[NR], tc2[NR], tc3[NR];
  {
    // rows r0+B-1 .. r0+B+4: the lower face of row r0 from the first five,
    // then the window
    T z[NR];
    int rr = qm_mod(r0 + B - 1, n2);
    const T* row = u + rr * rowP;
#pragma unroll
    for (int k = 0; k < NR; ++k) z[k] = row[i1[k]];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      rr = rr + 1 == n2 ? 0 : rr + 1;
      row = u + rr * rowP;
#pragma unroll
      for (int k = 0; k < NR; ++k) y[k][j] = row[i1[k]];
    }
    const int rs0 = qm_mod(r0, n2);
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      wc[k] = w_is_u ? T(0) : wv[rs0 * rowP + i1[k]];
      unc[k] = use_un ? unv[rs0 * rowP + i1[k]] : T(0);
      tc1[k] = src ? t1p[(long)rs0 * n1 + i1[k]] : T(0);
      tc2[k] = src2 ? t2p[(long)rs0 * n1 + i1[k]] : T(0);
      tc3[k] = src3 ? t3p[(long)rs0 * n1 + i1[k]] : T(0);
      s2m[k] = T(0);
      ka[k] = kb[k] = T(0);
      if (do_q2) {
        const T k0 = ql_kappa<R, T>(z[k], y[k][0], y[k][1]);
        ka[k] = ql_kappa<R, T>(y[k][0], y[k][1], y[k][2]);
        kb[k] = ql_kappa<R, T>(y[k][1], y[k][2], y[k][3]);
        s2m[k] = ql_face<R, P2, T>(z[k], y[k][0], y[k][1], y[k][2], y[k][3], k0, ka[k], kb[k]);
      }
    }
  }
  // periodic indices of the current row, the row after it and the window
  // row two rows ahead; the window row one ahead is already in flight
  int rcur = qm_mod(r0, n2);
  int rnx1 = rcur + 1 == n2 ? 0 : rcur + 1;
  int rnxw = qm_mod(r0 + B + 6, n2);
  T yn[NR];
  {
    const T* uw = u + qm_mod(r0 + B + 5, n2) * rowP;
#pragma unroll
    for (int k = 0; k < NR; ++k) yn[k] = __builtin_nontemporal_load(uw + i1[k]);
  }

  for (int it = 0; it < len; ++it) {
    // the next row's pointwise inputs one row ahead and the window row two
    // rows ahead, so loads stay in flight across the stencil work
    T ynn[NR], wn[NR], unn[NR], tn1[NR], tn2[NR], tn3[NR];
    {
      const T* uw = u + rnxw * rowP;
      const T* w1 = wv + rnx1 * rowP;
      const T* un1 = unv + rnx1 * rowP;
      const long tq = (long)rnx1 * n1;
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        ynn[k] = __builtin_nontemporal_load(uw + i1[k]);
        wn[k] = w_is_u ? T(0) : __builtin_nontemporal_load(w1 + i1[k]);
        unn[k] = use_un ? __builtin_nontemporal_load(un1 + i1[k]) : T(0);
        tn1[k] = src ? t1p[tq + i1[k]] : T(0);
        tn2[k] = src2 ? t2p[tq + i1[k]] : T(0);
        tn3[k] = src3 ? t3p[tq + i1[k]] : T(0);
      }
    }
    T d[NR], fc[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      fc[k] = P2 ? y[k][2] : y[k][1];
      d[k] = T(0);
    }
    T kc[NR];
    if (do_q2) {
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        kc[k] = ql_kappa<R, T>(y[k][2], y[k][3], y[k][4]);
        const T s2p = ql_face<R, P2, T>(y[k][0], y[k][1], y[k][2], y[k][3], y[k][4], ka[k],
                                        kb[k], kc[k]);
        d[k] = (C2 * s2p - C2 * s2m[k]) * tinv2;
        s2m[k] = s2p;
      }
    }
    if (do_q1) {
#pragma unroll
      for (int k = 0; k < NR; ++k) ql_put(fl, i1[k], live[k], n1, fc[k]);
      __builtin_amdgcn_wave_barrier();
      // P1: cells i-3, i-2, i-1, i+1; otherwise i-2, i-1, i+1, i+2
      T va[NR], vb[NR], vc[NR], vd[NR], kown[NR];
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        const T* line = fl + i1[k];
        va[k] = line[P1 ? -3 : -2];
        vb[k] = line[P1 ? -2 : -1];
        vc[k] = line[P1 ? -1 : 1];
        vd[k] = line[P1 ? 1 : 2];
        // neighbours i-1 and i+1 of the cell: (vc, vd) for P1, (vb, vc) otherwise
        kown[k] = P1 ? ql_kappa<R, T>(vc[k], fc[k], vd[k]) : ql_kappa<R, T>(vb[k], fc[k], vc[k]);
        ql_put(kl, i1[k], live[k], n1, kown[k]);
      }
      __builtin_amdgcn_wave_barrier();
      T s[NR];
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        if constexpr (P1) {
          // right face of cell i-1: window i-3..i+1, indicators of i-2, i-1, i
          T k2 = T(0), k1 = T(0);
          if constexpr (R == RC_WENO5) {
            k2 = kl[i1[k] - 2];
            k1 = kl[i1[k] - 1];
          }
          s[k] = ql_face<R, true, T>(va[k], vb[k], vc[k], fc[k], vd[k], k2, k1, kown[k]);
        } else {
          // left face of cell i: window i-2..i+2, indicators of i-1, i, i+1
          T k1 = T(0), kp = T(0);
          if constexpr (R == RC_WENO5) {
            k1 = kl[i1[k] - 1];
            kp = kl[i1[k] + 1];
          }
          s[k] = ql_face<R, false, T>(va[k], vb[k], fc[k], vc[k], vd[k], k1, kown[k], kp);
        }
        ql_put(sl, i1[k], live[k], n1, s[k]);
      }
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        const T su = sl[i1[k] + 1];  // lower face of cell i+1
        d[k] = -(C1 * su - C1 * s[k]) * tinv1 - d[k];
      }
      __builtin_amdgcn_wave_barrier();
    } else {
#pragma unroll
      for (int k = 0; k < NR; ++k) d[k] = -d[k];
    }
    // source, RK combination, store; slide the window
    T* orow = outv + rcur * rowP;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const T f = fc[k];
      T dk = d[k];
      if (src) {
        T fmb = tc1[k];
        if (src2) fmb = fmb * tc2[k];
        if (src3) fmb = fmb * tc3[k];
        dk = dk + (fmb - f) * tinvtau;
      }
      const T tv = (w_is_u ? f : wc[k]) + dk * tcdt;
      const T x = use_un ? ta * unc[k] + tb * tv : (plain ? tv : tb * tv);
      if (live[k]) __builtin_nontemporal_store(x, orow + i1[k]);
#pragma unroll
      for (int j = 0; j < 4; ++j) y[k][j] = y[k][j + 1];
      y[k][4] = yn[k];
      yn[k] = ynn[k];
      if (do_q2) {
        ka[k] = kb[k];
        kb[k] = kc[k];
      }
      wc[k] = wn[k];
      unc[k] = unn[k];
      tc1[k] = tn1[k];
      tc2[k] = tn2[k];
      tc3[k] = tn3[k];
    }
    rcur = rnx1;
    rnx1 = rnx1 + 1 == n2 ? 0 : rnx1 + 1;
    rnxw = rnxw + 1 == n2 ? 0 : rnxw + 1;
  }
}

// Blocks are numbered segment-major (all p-cells of one q2 segment, then the
// next segment) and spread over the eight XCDs in contiguous ranges, so the
// p-cells running together on one XCD share a segment and its rows of the
// Maxwellian tables stay in that XCD's L2.
template <typename T, int R, int NR, int MODE, int PD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void qline_kernel(QLParams q) {
  // wave-private q1 lines with three halo cells on each side: row values,
  // q1 indicator quantities and lower-face states
  __shared__ T fl[4][NR * 64 + 2 * QL_HALO];
  __shared__ T kl[4][NR * 64 + 2 * QL_HALO];
  __shared__ T sl[4][NR * 64 + 2 * QL_HALO];
  // the wave index is uniform across the wave: keep it (and all that derives
  // from it: p-cell, velocities, row pointers) in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int nb = gridDim.x;
  const int t = (nb % 8 == 0) ? (blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8 : blockIdx.x;
  const int npg = (q.P + 3) / 4;
  const int seg = t / npg;
  const int p = (t % npg) * 4 + wave;
  if (p >= q.P) return;  // no workgroup barriers below
  const int j1 = p % q.n_p1, j2 = (p / q.n_p1) % q.n_p2;
  const bool pos1 = centre(q.p1s, j1, q.dp1) > 0.0;  // wave-uniform
  const bool pos2 = centre(q.p2s, j2, q.dp2) > 0.0;
  T* f = fl[wave] + QL_HALO;
  T* k = kl[wave] + QL_HALO;
  T* s = sl[wave] + QL_HALO;
  if (pos1) {
    if (pos2) ql_march<T, R, NR, true, true, MODE, PD>(q, p, seg, lane, f, k, s);
    else ql_march<T, R, NR, true, false, MODE, PD>(q, p, seg, lane, f, k, s);
  } else {
    if (pos2) ql_march<T, R, NR, false, true, MODE, PD>(q, p, seg, lane, f, k, s);
    else ql_march<T, R, NR, false, false, MODE, PD>(q, p, seg, lane, f, k, s);
  }
}

}  // namespace bx
