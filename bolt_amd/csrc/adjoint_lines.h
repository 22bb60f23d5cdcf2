// Stage adjoint in the q1-line layout [q2][p][q1] for WENO5 (FP64), with every
// WENO5 indicator and face evaluated once per cell and axis (reverse mode),
// instead of once per cell of each face's stencil as in adj_stage_kernel.
//
// For one axis with upwind face states s_k (the lower face of cell k, built
// from the window of cells b..b+4, b = k-3 for C > 0 and k-2 otherwise, and
// the indicators beta of cells b+1..b+3; weno5.py:57-98), the adjoint of the
// advective term of L seen through mu is
//   sum_k g_k ds_k/du,   g_k = (C/dq) (mu_k - mu_{k-1})
// (df_dt_fvm.py:56-85).  With weights W_m = c_m / (16 beta_m^2) (c = (1, 10,
// 5) for a right face, (5, 10, 1) for a left one), D = sum W_m, substencil
// values u_m and s = sum W_m u_m / D:
//   A_m = g W_m / D                      (adjoint of u_m)
//   h_m = -2 A_m (u_m - s) / beta_m      (adjoint of beta_m)
// and cell j collects sum_m A_m coef(m, j - b) over the five faces whose
// window holds it, plus H_c dbeta_c/du_j for the indicators c = j-1, j, j+1,
// H_c being the sum of the h that name c.
//
// adj_line_kernel: a wave owns one p-cell and the whole q1 line (NR cells
// per lane, lane l holding q1 = l + 64 k) and marches along q2 over a segment
// of rows, as qline_kernel does forward:
//  * q2: at step t it loads row t, forms the indicator of cell t-1 and the
//    face whose window is rows t-4..t, adds that face's A into the pending
//    adjoints of rows t-4..t and its h into the indicator adjoints of cells
//    t-3..t-1; the indicator adjoint of cell t-3 is then complete and goes
//    into rows t-4..t-2, and row t-4 is complete;
//  * q1: the completed row's u and mu go through wave-private LDS lines
//    (indicators, then faces, then the gather, which sums each indicator's
//    adjoint from the three faces that name it), so no workgroup barrier is
//    needed.
#pragma once
#include "adjoint.h"

namespace bx {

constexpr int AL_HALO = 4;   // periodic halo cells of the LDS lines
constexpr int AL_LINES = 10; // U, M, R, IB, A1..A3, h1..h3

// Substencil values of a right (or left) face window y[0..4].
template <bool RIGHT>
__device__ __forceinline__ void al_substencils(const double* y, double u[3]) {
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    double acc = 0.0;
#pragma unroll
    for (int e = 0; e < 5; ++e)
      if (weno_sub_coef<RIGHT>(m + 1, e) != 0.0) acc += weno_sub_coef<RIGHT>(m + 1, e) * y[e];
    u[m] = acc;
  }
}

// Reverse mode of one face: A (adjoints of the substencil values) and h
// (adjoints of the three indicators) from g, the window y[0..4], the
// indicator quantities R = 1/(16 beta^2) and IB = 1/beta of cells b+1..b+3.
template <bool RIGHT>
__device__ __forceinline__ void al_face(const double* y, const double* R, const double* IB,
                                        double g, double A[3], double h[3]) {
  constexpr double c1 = RIGHT ? 1.0 : 5.0, c3 = RIGHT ? 5.0 : 1.0;
  double u[3];
  al_substencils<RIGHT>(y, u);
  const double W1 = c1 * R[0], W2 = 10.0 * R[1], W3 = c3 * R[2];
  const double rD = fast_rcp(W1 + W2 + W3);
  const double sv = (W1 * u[0] + W2 * u[1] + W3 * u[2]) * rD;
  const double gr = g * rD;
  A[0] = gr * W1;
  A[1] = gr * W2;
  A[2] = gr * W3;
#pragma unroll
  for (int m = 0; m < 3; ++m) h[m] = -2.0 * A[m] * (u[m] - sv) * IB[m];
}

// Add a face's A into the adjoints of its window cells aw[0..4].
template <bool RIGHT>
__device__ __forceinline__ void al_scatter(const double A[3], double* aw) {
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    double x = 0.0;
#pragma unroll
    for (int m = 0; m < 3; ++m)
      if (weno_sub_coef<RIGHT>(m + 1, e) != 0.0) x += A[m] * weno_sub_coef<RIGHT>(m + 1, e);
    aw[e] += x;
  }
}

// Store v (NR cells of the lane) into a halo-padded LDS line with its
// periodic copies (n1 >= 64 > 2 * AL_HALO).
template <int NR>
__device__ __forceinline__ void al_put(double* line, const double* v, int lane, int n1) {
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const int i = lane + 64 * k;
    line[i] = v[k];
    if (i < AL_HALO) line[i + n1] = v[k];
    if (i >= n1 - AL_HALO) line[i - n1] = v[k];
  }
}

// Adjoint of the q1 advection for one completed row: u and mu of the row
// (NR cells per lane) in, the q1 part of J^T mu for the lane's cells out.
template <int NR, bool P1>
__device__ __forceinline__ void al_q1_row(const double* ur, const double* mr, double cdq1, int lane,
                                          int n1, double* lds, double* out) {
  const int LS = n1 + 2 * AL_HALO;
  double* U = lds + AL_HALO;
  double* M = U + LS;
  double* Rl = M + LS;
  double* IBl = Rl + LS;
  double* Al = IBl + LS;   // three lines
  double* hl = Al + 3 * LS;  // three lines
  al_put<NR>(U, ur, lane, n1);
  al_put<NR>(M, mr, lane, n1);
  __builtin_amdgcn_wave_barrier();
  // indicators of the lane's cells
  double v[NR], w[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const int c = lane + 64 * k;
    const double B = weno_beta<double>(U[c - 1], U[c], U[c + 1]);
    // 1/(16 beta^2) from 1/beta: one reciprocal per indicator
    w[k] = fast_rcp(B);
    v[k] = 0.0625 * (w[k] * w[k]);
  }
  al_put<NR>(Rl, v, lane, n1);
  al_put<NR>(IBl, w, lane, n1);
  __builtin_amdgcn_wave_barrier();
  // the lower face of each of the lane's cells
  double A[NR][3], h[NR][3];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const int c = lane + 64 * k;
    const int b = P1 ? c - 3 : c - 2;
    double y[5], R[3], IB[3];
#pragma unroll
    for (int e = 0; e < 5; ++e) y[e] = U[b + e];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      R[m] = Rl[b + 1 + m];
      IB[m] = IBl[b + 1 + m];
    }
    al_face<P1>(y, R, IB, cdq1 * (M[c] - M[c - 1]), A[k], h[k]);
  }
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    double a[NR], x[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      a[k] = A[k][m];
      x[k] = h[k][m];
    }
    al_put<NR>(Al + m * LS, a, lane, n1);
    al_put<NR>(hl + m * LS, x, lane, n1);
  }
  __builtin_amdgcn_wave_barrier();
  // gather: substencil adjoints of the five faces whose window holds the
  // cell, then the indicator adjoints of cells c-1, c, c+1
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const int c = lane + 64 * k;
    double acc = 0.0;
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      const int f = P1 ? c + 3 - e : c + 2 - e;  // the face whose window has c at entry e
#pragma unroll
      for (int m = 0; m < 3; ++m)
        if (weno_sub_coef<P1>(m + 1, e) != 0.0) acc += Al[m * LS + f] * weno_sub_coef<P1>(m + 1, e);
    }
    // indicator adjoint of cell x: face k names x as indicator m when
    // k = x + 2 - m (C > 0) or k = x + 1 - m (C <= 0)
    auto Hc = [&](int x) {
      const int k0 = P1 ? x + 2 : x + 1;
      return (hl[k0] + hl[LS + k0 - 1]) + hl[2 * LS + k0 - 2];
    };
    acc += Hc(c + 1) * beta_da(U[c], U[c + 1], U[c + 2]);
    acc += Hc(c) * beta_db(U[c - 1], U[c], U[c + 1]);
    acc += Hc(c - 1) * beta_dc(U[c - 2], U[c - 1], U[c]);
    out[k] = acc;
  }
  __builtin_amdgcn_wave_barrier();
}

// One wave's march over the q2 segment ``seg`` of p-cell p (P1: C_q1 > 0,
// P2: C_q2 > 0, wave-uniform).
template <int NR, bool P1, bool P2>
__device__ void al_march(const AdjParams& ap, int p, int seg, int nseg, int lane, double* lds) {
  const int n1 = ap.n1, n2 = ap.n2, P = ap.P;
  const int S = (n2 + nseg - 1) / nseg;
  const int r0 = seg * S, r1 = min(n2, r0 + S);
  if (r0 >= n2) return;
  int j1, j2;
  const int j23 = fdivmod(p, ap.n_p1, 1.0f / float(ap.n_p1), j1);
  const int j3 = fdivmod(j23, ap.n_p2, 1.0f / float(ap.n_p2), j2);
  const double v1 = centre(ap.p1s, j1, ap.dp1), v2 = centre(ap.p2s, j2, ap.dp2);
  const double v3 = centre(ap.p3s, j3, ap.dp3);
  const double e3 = 0.5 * ((v1 * v1 + v2 * v2) + v3 * v3);
  const double cdq1 = v1 * ap.inv_dq1;
  const double cdq2 = n2 > 1 ? v2 * ap.inv_dq2 : 0.0;
  const long rs = (long)n1 * P, col = (long)p * n1;
  const long Qn = (long)n1 * n2;
  // windows, [row][cell]: u and mu of rows t-4..t, indicator quantities of
  // cells t-3..t-1, their adjoints, and the pending adjoints of rows t-4..t
  double uw[5][NR], mw[5][NR], Rw[3][NR], IBw[3][NR], Hw[3][NR], aw[5][NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
#pragma unroll
    for (int r = 0; r < 5; ++r) uw[r][k] = mw[r][k] = aw[r][k] = 0.0;
#pragma unroll
    for (int r = 0; r < 3; ++r) Rw[r][k] = IBw[r][k] = Hw[r][k] = 0.0;
  }
  for (int t = r0 - 4; t <= r1 + 3; ++t) {
    int row = t % n2;
    if (row < 0) row += n2;
    const double* ur = ap.u + row * rs + col;
    const double* mr = ap.lam + row * rs + col;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        uw[r][k] = uw[r + 1][k];
        mw[r][k] = mw[r + 1][k];
        aw[r][k] = aw[r + 1][k];
      }
      uw[4][k] = ur[lane + 64 * k];
      mw[4][k] = mr[lane + 64 * k];
      aw[4][k] = 0.0;
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        Rw[r][k] = Rw[r + 1][k];
        IBw[r][k] = IBw[r + 1][k];
        Hw[r][k] = Hw[r + 1][k];
      }
      // indicator of cell t-1
      const double B = weno_beta<double>(uw[2][k], uw[3][k], uw[4][k]);
      IBw[2][k] = fast_rcp(B);
      Rw[2][k] = 0.0625 * (IBw[2][k] * IBw[2][k]);
      Hw[2][k] = 0.0;
    }
    if (t >= r0) {
      // the face with window rows t-4..t: the lower face of row t-1 (C > 0)
      // or of row t-2
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        const double g = P2 ? cdq2 * (mw[3][k] - mw[2][k]) : cdq2 * (mw[2][k] - mw[1][k]);
        double y[5], R[3], IB[3], A[3], h[3], a[5];
#pragma unroll
        for (int r = 0; r < 5; ++r) {
          y[r] = uw[r][k];
          a[r] = aw[r][k];
        }
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          R[m] = Rw[m][k];
          IB[m] = IBw[m][k];
        }
        al_face<P2>(y, R, IB, g, A, h);
        al_scatter<P2>(A, a);
#pragma unroll
        for (int r = 0; r < 5; ++r) aw[r][k] = a[r];
#pragma unroll
        for (int m = 0; m < 3; ++m) Hw[m][k] += h[m];
      }
    }
    if (t >= r0 + 2) {
      // the indicator adjoint of cell t-3 is complete: rows t-4, t-3, t-2
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        const double H = Hw[0][k];
        aw[0][k] += H * beta_da(uw[0][k], uw[1][k], uw[2][k]);
        aw[1][k] += H * beta_db(uw[0][k], uw[1][k], uw[2][k]);
        aw[2][k] += H * beta_dc(uw[0][k], uw[1][k], uw[2][k]);
      }
    }
    if (t >= r0 + 4) {
      // row j = t-4 is complete
      const int j = t - 4;
      double q1[NR], ujr[NR], mjr[NR];
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        ujr[k] = uw[0][k];
        mjr[k] = mw[0][k];
      }
      al_q1_row<NR, P1>(ujr, mjr, cdq1, lane, n1, lds, q1);
      const long base = j * rs + col;
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        const int i1 = lane + 64 * k;
        const double mu = mjr[k];
        double acc = q1[k] + aw[0][k];
        if (ap.source) {
          const double* g = ap.gq + (long)j * n1 + i1;
          acc += ap.inv_tau * (ap.measure * (g[0] + g[Qn] * v1 + g[2 * Qn] * v2 + g[3 * Qn] * v3 +
                                             g[4 * Qn] * e3) - mu);
        }
        double x = ap.b * mu + ap.c * acc;
        if (ap.base) x += ap.a * ap.base[base + i1];
        if (ap.base2) x += ap.a2 * ap.base2[base + i1];
        ap.out[base + i1] = x;
      }
    }
  }
}

#ifndef BX_NO_KERNELS
// Blocks are numbered segment-major and spread over the eight XCDs in
// contiguous ranges (as for qline_kernel); wave w of a block takes p-cell
// 4 * (block's p-group) + w.
// Four waves per SIMD (at most 128 VGPRs): the march waits on LDS and
// global loads, and the fourth wave hides more of that than the registers
// it costs (measured: 76.6 -> 70 us per launch at 64^2 x 32^2).
template <int NR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void adj_line_kernel(AdjParams ap, int nseg) {
  __shared__ double lds[4][AL_LINES * (NR * 64 + 2 * AL_HALO)];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int nb = gridDim.x;
  const int t = (nb % 8 == 0) ? (blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8 : blockIdx.x;
  const int npg = (ap.P + 3) / 4;
  const int seg = t / npg;
  const int p = (t % npg) * 4 + wave;
  if (p >= ap.P) return;  // no workgroup barriers below
  int j1, j2;
  fdivmod(p, ap.n_p1, 1.0f / float(ap.n_p1), j1);
  const int j23 = p / ap.n_p1;
  j2 = j23 % ap.n_p2;
  const bool pos1 = centre(ap.p1s, j1, ap.dp1) > 0.0;
  const bool pos2 = centre(ap.p2s, j2, ap.dp2) > 0.0;
  double* l = lds[wave];
  if (pos1) {
    if (pos2) al_march<NR, true, true>(ap, p, seg, nseg, lane, l);
    else al_march<NR, true, false>(ap, p, seg, nseg, lane, l);
  } else {
    if (pos2) al_march<NR, false, true>(ap, p, seg, nseg, lane, l);
    else al_march<NR, false, false>(ap, p, seg, nseg, lane, l);
  }
}
#endif

}  // namespace bx
