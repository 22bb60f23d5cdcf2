// The nonlinear FVM stage in the q1-line layout [q2][p][q1] (q1 fastest), for
// 2-D or 1-D q-space with fields off -- the layout the solver picks when the
// q1 extent is a multiple of 64.
//
// qline_kernel: a wave owns one p-cell and the whole periodic q1 line (NR
// cells per lane, lane l holding q1 = l + 64 k), and marches along q2 over a
// segment of L rows:
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
  int n1, n2, P, L, nseg, n_p1, n_p2, n_p3;
  const void* u;
  const void* w;
  const void* un;
  void* out;
  double a, b, cdt, inv_dq1, inv_dq2, inv_tau;
  int source, p_dim;
  double p1s, dp1, p2s, dp2, p3s, dp3;
  const void* tab;  // [n_p1 + n_p2 + n_p3][Q] of T: prefactor*e1, e2, e3
};

// Upwind face state at the lower face of the centre cell a3 of a[k] =
// f[i-3+k] along an axis, with indicators bm = beta(i-2), b0 = beta(i-1),
// bc = beta(i), bp = beta(i+1): the right face of cell i-1 (pos) or the left
// face of cell i.
template <int R, typename T>
__device__ inline T ql_lower_face(T a0, T a1, T a2, T a3, T a4, T a5, T bm, T b0, T bc, T bp,
                                  bool pos) {
  if constexpr (R == RC_WENO5) {
    return pos ? weno_face_betas(a0, a1, a2, a3, a4, bm, b0, bc, true)
               : weno_face_betas(a1, a2, a3, a4, a5, b0, bc, bp, false);
  } else {
    return pos ? face_value<R>(a0, a1, a2, a3, a4, true) : face_value<R>(a1, a2, a3, a4, a5, false);
  }
}

template <int R, typename T>
__device__ inline T ql_beta(T a, T b, T c) {
  if constexpr (R == RC_WENO5) return weno_beta<T>(a, b, c);
  else return T(0);
}

template <typename T, int R, int NR>
__global__ __launch_bounds__(256) void qline_kernel(QLParams q) {
  __shared__ T fl[4][NR * 64];  // wave-private q1 lines: row values,
  __shared__ T bl[4][NR * 64];  // q1 indicators,
  __shared__ T sl[4][NR * 64];  // q1 lower-face states
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nb = gridDim.x;
  const int t = (nb % 8 == 0) ? (blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8 : blockIdx.x;
  const int seg = t % q.nseg;
  const int p = (t / q.nseg) * 4 + wave;
  if (p >= q.P) return;  // no workgroup barriers below
  const int n1 = q.n1, n2 = q.n2;
  const int j1 = p % q.n_p1, j23 = p / q.n_p1, j2 = j23 % q.n_p2, j3 = j23 / q.n_p2;
  const double v1 = centre(q.p1s, j1, q.dp1);
  const double v2 = centre(q.p2s, j2, q.dp2);
  const bool pos1 = v1 > 0.0, pos2 = v2 > 0.0;  // wave-uniform
  const T C1 = T(v1), C2 = T(v2);
  const long rowP = (long)q.P * n1;  // elements per q2 row
  const int r0 = seg * q.L;
  // per-wave base pointers of this p-cell; rows are added as uniform offsets
  const T* u = (const T*)q.u + (long)p * n1;
  const T* wv = (const T*)q.w + (long)p * n1;
  const T* unv = (const T*)q.un + (long)p * n1;
  T* outv = (T*)q.out + (long)p * n1;
  const long Q = (long)n1 * n2;
  const T* t1p = (const T*)q.tab + (long)j1 * Q;
  const T* t2p = (const T*)q.tab + (long)(q.n_p1 + j2) * Q;
  const T* t3p = (const T*)q.tab + (long)(q.n_p1 + q.n_p2 + j3) * Q;
  const bool use_un = q.a != 0.0, w_is_u = q.w == q.u;
  const T ta = T(q.a), tb = T(q.b), tcdt = T(q.cdt);
  const T tinv1 = T(q.inv_dq1), tinv2 = T(q.inv_dq2), tinvtau = T(q.inv_tau);
  const bool do_q1 = n1 > 1, do_q2 = n2 > 1;
  const bool src = q.source != 0, src2 = src && q.p_dim >= 2, src3 = src && q.p_dim >= 3;

  // lane cells and their periodic q1 neighbours (LDS indices), fixed for the kernel
  int i1[NR], nm3[NR], nm2[NR], nm1[NR], np1[NR], np2[NR];
  bool live[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    i1[k] = lane + 64 * k;
    live[k] = i1[k] < n1;
    if (!live[k]) i1[k] = 0;
    nm3[k] = qm_mod(i1[k] - 3, n1);
    nm2[k] = qm_mod(i1[k] - 2, n1);
    nm1[k] = qm_mod(i1[k] - 1, n1);
    np1[k] = qm_mod(i1[k] + 1, n1);
    np2[k] = qm_mod(i1[k] + 2, n1);
  }

  T y[NR][7], bq[NR][4], s2m[NR], wc[NR], unc[NR], tc1[NR], tc2[NR], tc3[NR];
  {
    int rr = qm_mod(r0 - 3, n2);
#pragma unroll
    for (int j = 0; j < 7; ++j) {
      const T* row = u + rr * rowP;
#pragma unroll
      for (int k = 0; k < NR; ++k) y[k][j] = row[i1[k]];
      rr = rr + 1 == n2 ? 0 : rr + 1;
    }
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
#pragma unroll
    for (int j = 0; j < 4; ++j) bq[k][j] = T(0);
    if (do_q2) {
      // indicators of rows r0-2 .. r0+1 and the state at face r0-1/2
      const T bm2 = ql_beta<R, T>(y[k][0], y[k][1], y[k][2]);
      bq[k][0] = ql_beta<R, T>(y[k][1], y[k][2], y[k][3]);
      bq[k][1] = ql_beta<R, T>(y[k][2], y[k][3], y[k][4]);
      bq[k][2] = ql_beta<R, T>(y[k][3], y[k][4], y[k][5]);
      s2m[k] = ql_lower_face<R, T>(y[k][0], y[k][1], y[k][2], y[k][3], y[k][4], y[k][5], bm2,
                                   bq[k][0], bq[k][1], bq[k][2], pos2);
    }
  }
  // periodic row indices of r (current), r+1 and r+4
  int rcur = rs0;
  int rnx1 = rcur + 1 == n2 ? 0 : rcur + 1;
  int rnx4 = qm_mod(r0 + 4, n2);

  for (int it = 0; it < q.L; ++it) {
    // prefetch for the next row
    T yn[NR], wn[NR], unn[NR], tn1[NR], tn2[NR], tn3[NR];
    {
      const T* u4 = u + rnx4 * rowP;
      const T* w1 = wv + rnx1 * rowP;
      const T* un1 = unv + rnx1 * rowP;
      const long tq = (long)rnx1 * n1;
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        yn[k] = u4[i1[k]];
        wn[k] = w_is_u ? T(0) : w1[i1[k]];
        unn[k] = use_un ? un1[i1[k]] : T(0);
        tn1[k] = src ? t1p[tq + i1[k]] : T(0);
        tn2[k] = src2 ? t2p[tq + i1[k]] : T(0);
        tn3[k] = src3 ? t3p[tq + i1[k]] : T(0);
      }
    }
    T d[NR];
    // q2 direction
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      d[k] = T(0);
      if (do_q2) {
        bq[k][3] = ql_beta<R, T>(y[k][4], y[k][5], y[k][6]);
        // state at face r+1/2 = lower face of row r+1: rows r-2..r+3
        const T s2p = ql_lower_face<R, T>(y[k][1], y[k][2], y[k][3], y[k][4], y[k][5], y[k][6],
                                          bq[k][0], bq[k][1], bq[k][2], bq[k][3], pos2);
        d[k] = (C2 * s2p - C2 * s2m[k]) * tinv2;
        s2m[k] = s2p;
      }
    }
    // q1 direction through the wave's LDS line
    if (do_q1) {
#pragma unroll
      for (int k = 0; k < NR; ++k)
        if (live[k]) fl[wave][i1[k]] = y[k][3];
      __builtin_amdgcn_wave_barrier();
      T fm3[NR], fm2[NR], fm1[NR], fp1[NR], fp2[NR], b1[NR];
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        fm3[k] = fl[wave][nm3[k]];
        fm2[k] = fl[wave][nm2[k]];
        fm1[k] = fl[wave][nm1[k]];
        fp1[k] = fl[wave][np1[k]];
        fp2[k] = fl[wave][np2[k]];
        b1[k] = ql_beta<R, T>(fm1[k], y[k][3], fp1[k]);
        if (live[k]) bl[wave][i1[k]] = b1[k];
      }
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        T bm = T(0), b0 = T(0), bp = T(0);
        if constexpr (R == RC_WENO5) {
          bm = bl[wave][nm2[k]];
          b0 = bl[wave][nm1[k]];
          bp = bl[wave][np1[k]];
        }
        const T s = ql_lower_face<R, T>(fm3[k], fm2[k], fm1[k], y[k][3], fp1[k], fp2[k], bm, b0,
                                        b1[k], bp, pos1);
        if (live[k]) sl[wave][i1[k]] = s;
        fm3[k] = s;  // keep the lower-face state
      }
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        const T su = sl[wave][np1[k]];
        d[k] = -(C1 * su - C1 * fm3[k]) * tinv1 - d[k];
      }
      __builtin_amdgcn_wave_barrier();
    } else {
#pragma unroll
      for (int k = 0; k < NR; ++k) d[k] = -d[k];
    }
    // source, RK combination, store
    T* orow = outv + rcur * rowP;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const T f = y[k][3];
      T dk = d[k];
      if (src) {
        T fmb = tc1[k];
        if (src2) fmb = fmb * tc2[k];
        if (src3) fmb = fmb * tc3[k];
        dk = dk + (fmb - f) * tinvtau;
      }
      const T tv = (w_is_u ? f : wc[k]) + dk * tcdt;
      const T x = use_un ? ta * unc[k] + tb * tv : (q.b == 1.0 ? tv : tb * tv);
      if (live[k]) orow[i1[k]] = x;
      // slide
#pragma unroll
      for (int j = 0; j < 6; ++j) y[k][j] = y[k][j + 1];
      y[k][6] = yn[k];
      bq[k][0] = bq[k][1];
      bq[k][1] = bq[k][2];
      bq[k][2] = bq[k][3];
      wc[k] = wn[k];
      unc[k] = unn[k];
      tc1[k] = tn1[k];
      tc2[k] = tn2[k];
      tc3[k] = tn3[k];
    }
    rcur = rnx1;
    rnx1 = rnx1 + 1 == n2 ? 0 : rnx1 + 1;
    rnx4 = rnx4 + 1 == n2 ? 0 : rnx4 + 1;
  }
}

// Moments of f in the q1-line layout: workgroup (q2 row, 64-wide q1 block,
// p chunk); wave w sums a contiguous quarter of the chunk's p range for its
// 64 q1 cells (loads issued eight at a time), and the four waves are added in
// a fixed order -> part[chunk][5][Q].
template <typename T>
__global__ __launch_bounds__(256) void mline_kernel(const T* f, int n1, int n2, int P, int nchunk,
                                                    int n_p1, int n_p2, double p1s, double dp1,
                                                    double p2s, double dp2, double p3s, double dp3,
                                                    double* part) {
  __shared__ double red[5][4][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nblk = (n1 + 63) / 64;
  int b = blockIdx.x;
  const int blk = b % nblk;
  b /= nblk;
  const int i2 = b % n2;
  const int chunk = b / n2;
  const int i1 = blk * 64 + lane;
  const bool live = i1 < n1;
  const int per = (P + 4 * nchunk - 1) / (4 * nchunk);
  const int pa = min(P, (chunk * 4 + wave) * per), pe = min(P, pa + per);
  const T* col = f + (long)i2 * P * n1 + (live ? i1 : 0);
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0;
  int j1 = pa % n_p1, j2 = (pa / n_p1) % n_p2, j3 = pa / (n_p1 * n_p2);
  double v2 = centre(p2s, j2, dp2), v3 = centre(p3s, j3, dp3);
  for (int p = pa; p < pe; p += 8) {
    double x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = (p + k < pe && live) ? double(col[(long)(p + k) * n1]) : 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (p + k < pe) {
        const double v1 = centre(p1s, j1, dp1);
        const double e = 0.5 * ((v1 * v1 + v2 * v2) + v3 * v3);
        a0 += x[k];
        a1 += x[k] * v1;
        a2 += x[k] * v2;
        a3 += x[k] * v3;
        a4 += x[k] * e;
        if (++j1 == n_p1) {
          j1 = 0;
          if (++j2 == n_p2) {
            j2 = 0;
            ++j3;
            v3 = centre(p3s, j3, dp3);
          }
          v2 = centre(p2s, j2, dp2);
        }
      }
    }
  }
  red[0][wave][lane] = a0;
  red[1][wave][lane] = a1;
  red[2][wave][lane] = a2;
  red[3][wave][lane] = a3;
  red[4][wave][lane] = a4;
  __syncthreads();
  if (wave == 0 && live) {
    const long Q = (long)n1 * n2;
    const long qc = (long)i2 * n1 + i1;
    for (int k = 0; k < 5; ++k)
      part[((long)chunk * 5 + k) * Q + qc] =
          ((red[k][0][lane] + red[k][1][lane]) + red[k][2][lane]) + red[k][3][lane];
  }
}

// Sum the chunk partials -> moment block (times dp3*dp2*dp1).
__global__ void ml_finalize_kernel(const double* part, int nchunk, int Q, double measure,
                                   double* mom, double mass, double kB, int p_dim) {
  const int qc = blockIdx.x * blockDim.x + threadIdx.x;
  if (qc >= Q) return;
  double m[5];
  for (int k = 0; k < 5; ++k) {
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += part[((long)c * 5 + k) * Q + qc];
    m[k] = s * measure;
  }
  write_moment_block(mom, Q, qc, m, mass, kB, p_dim);
}

// Maxwellian tables of every q-cell from its moment block
// (collision_operator.py:52-65): tab[j1][q] = prefactor * exp(-m (v1-u1)^2/2kT),
// tab[n_p1 + j2][q] = exp(-m (v2-u2)^2/2kT), tab[n_p1 + n_p2 + j3][q] likewise.
template <typename T>
__global__ void bgk_table_kernel(const double* mom, int Q, int n_p1, int n_p2, int n_p3,
                                 double p1s, double dp1, double p2s, double dp2, double p3s,
                                 double dp3, double mass, double kB, T* tab) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nj = n_p1 + n_p2 + n_p3;
  if (t >= (long)nj * Q) return;
  const int j = (int)(t / Q), qc = (int)(t % Q);
  const double* b = mom + MOM_BGK * Q + qc;
  const double Tq = b[4 * Q];
  double val;
  if (j < n_p1) val = b[0] * maxwell_factor(centre(p1s, j, dp1), b[Q], Tq, mass, kB);
  else if (j < n_p1 + n_p2)
    val = maxwell_factor(centre(p2s, j - n_p1, dp2), b[2 * Q], Tq, mass, kB);
  else val = maxwell_factor(centre(p3s, j - n_p1 - n_p2, dp3), b[3 * Q], Tq, mass, kB);
  tab[t] = T(val);
}

}  // namespace bx
