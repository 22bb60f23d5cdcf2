// Adjoint (vector-Jacobian product) of one FVM stage, FP64, fields off.
//
// The stage is out = a*un + b*(w + cdt*L(u)) with
//   L(u) = -(F1(i+1/2) - F1(i-1/2))/dq1 - (F2(j+1/2) - F2(j-1/2))/dq2 + (f_MB(u) - u)/tau
// (df_dt_fvm.py:56-97, collision_operator.py:70-128), where each face flux is
// C times the upwind face state (riemann.py:46-49; LLF has the same
// derivative for a linear flux).  Given mu (the adjoint of the stage output
// seen through L), adj_stage_kernel writes
//   out = a*base + a2*base2 + b*mu + c*J(u)^T mu,
// with J = dL/du, and adj_bgk_reduce_kernel the per-q-cell pieces of the
// collision term's Jacobian and c * sum_p mu dL/dtau.
//
// Advection: sum_i mu_i L_i = sum_k F_k (mu_k - mu_{k-1}) / dq over the lower
// faces k, so the face state s_k has adjoint C (mu_k - mu_{k-1}) / dq, and the
// gather for cell j visits the five faces whose stencils contain j.  Each
// face's derivative with respect to cell j is the reference reconstruction
// (fvm_math.h, the same templates the forward kernels use) evaluated on
// one-direction dual numbers.
//
// BGK: f_MB = n (m/2 pi k T)^(d/2) prod_{k<=d} exp(-m (v_k-u_k)^2 / 2kT) with
// n = M0, u_k = M_k/(n+eps), T = (2 E m - n m |u|^2)/(d (n+eps)) + eps
// (collision_operator.py:102-116) and the raw moments M = sum_p f w_p dV.  The
// reduction over p of mu f_MB, mu f_MB (v-u), mu f_MB |v-u|^2 gives the
// derivatives with respect to n, u, T, which the chain rule through
// u(M), T(M) turns into one weight per moment; the pointwise pass then adds
// dV (g0 + g1 v1 + g2 v2 + g3 v3 + g4 |v|^2/2) / tau - mu / tau.
#pragma once
#include "../../include/boltx.h"
#include "fvm_math.h"

namespace bx {

// Forward-mode dual number: value and one directional derivative.
struct Dual {
  double v, d;
  __host__ __device__ Dual() : v(0.0), d(0.0) {}
  __host__ __device__ Dual(double x) : v(x), d(0.0) {}
  __host__ __device__ Dual(double x, double dx) : v(x), d(dx) {}
};
__host__ __device__ inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__host__ __device__ inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__host__ __device__ inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__host__ __device__ inline Dual operator*(Dual a, Dual b) {
  return Dual(a.v * b.v, a.d * b.v + a.v * b.d);
}
__host__ __device__ inline Dual operator/(Dual a, Dual b) {
  const double r = fast_rcp(b.v);
  return Dual(a.v * r, (a.d - a.v * r * b.d) * r);
}
__host__ __device__ inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
__host__ __device__ inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
__host__ __device__ inline bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
__host__ __device__ inline bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
__host__ __device__ inline Dual aabs(Dual a) { return a.v < 0.0 ? -a : a; }
__host__ __device__ inline Dual fast_rcp(Dual a) {
  const double r = fast_rcp(a.v);
  return Dual(r, -a.d * r * r);
}
// weno5.py:58-71 (expanded indicator plus the eps floor) on duals
template <>
__host__ __device__ inline Dual weno_beta<Dual>(Dual a, Dual b, Dual c) {
  const Dual s = Dual(4.0 / 3.0) * a * a - Dual(19.0 / 3.0) * a * b + Dual(25.0 / 3.0) * b * b +
                 Dual(11.0 / 3.0) * a * c - Dual(31.0 / 3.0) * b * c + Dual(10.0 / 3.0) * c * c;
  return s + Dual(1e-17) * (Dual(1.0) + aabs(a) + aabs(b) + aabs(c));
}

// Derivative of the face state of the centre cell w[s+2] of the window
// w[s..s+4] (its right face if ``right``) with respect to the window entry
// w[4].
template <int R>
__host__ __device__ inline double face_partial(const double* w, int s, bool right) {
  Dual y[5];
#pragma unroll
  for (int r = 0; r < 5; ++r) y[r] = Dual(w[s + r], s + r == 4 ? 1.0 : 0.0);
  return face_value<R, Dual>(y[0], y[1], y[2], y[3], y[4], right).d;
}

// WENO5 (weno5.py:58-96 as the forward kernels evaluate it, FP64): the
// derivatives of the five upwind face states whose windows hold the gather
// cell w[4] with respect to it, analytically and with the shared parts
// computed once.  Face f of the gather has window w[f..f+4] and indicators
// b_k = beta(w[f+k-1..f+k+1]) of the seven windows starting at w[0..6]; with
// r_k = 1 / (16 beta_k^2) its weights are W = (c1 r_f, 10 r_{f+1}, c3 r_{f+2})
// ((c1, c3) = (1, 5) for a right face, (5, 1) for a left one) and its state is
// s = sum W_k u_k / sum W_k, so
//   ds/dw4 = (sum_k dW_k (u_k - s) + W_k du_k) / sum W_k,
//   dW_k = -2 W_k dbeta_k / beta_k,
// where only the indicators of the windows starting at w[2], w[3], w[4]
// depend on w[4].
__host__ __device__ inline double beta_da(double a, double b, double c) {
  return (8.0 / 3.0) * a - (19.0 / 3.0) * b + (11.0 / 3.0) * c + (a < 0.0 ? -1e-17 : 1e-17);
}
__host__ __device__ inline double beta_db(double a, double b, double c) {
  return -(19.0 / 3.0) * a + (50.0 / 3.0) * b - (31.0 / 3.0) * c + (b < 0.0 ? -1e-17 : 1e-17);
}
__host__ __device__ inline double beta_dc(double a, double b, double c) {
  return (11.0 / 3.0) * a - (31.0 / 3.0) * b + (20.0 / 3.0) * c + (c < 0.0 ? -1e-17 : 1e-17);
}

// Coefficient of window entry e (0..4) in substencil k (1..3) of a right or
// left face (weno5.py:86-92).
template <bool RIGHT>
__host__ __device__ constexpr double weno_sub_coef(int k, int e) {
  if (k == 1) return e == 0 ? (RIGHT ? 0.375 : -0.125) : e == 1 ? (RIGHT ? -1.25 : 0.75)
                   : e == 2 ? (RIGHT ? 1.875 : 0.375) : 0.0;
  if (k == 2) return e == 1 ? (RIGHT ? -0.125 : 0.375) : e == 2 ? 0.75
                   : e == 3 ? (RIGHT ? 0.375 : -0.125) : 0.0;
  return e == 2 ? (RIGHT ? 0.375 : 1.875) : e == 3 ? (RIGHT ? 0.75 : -1.25)
         : e == 4 ? (RIGHT ? -0.125 : 0.375) : 0.0;
}

// One face of the gather (window w[F..F+4], gather cell at entry E = 4 - F),
// with only the terms that do not vanish: indicator partials of the windows
// starting at w[2..4] (G[c] = -2 dbeta_c / beta_c, so dW = W G), substencil
// coefficients of entry E.
template <bool RIGHT, int F>
__host__ __device__ inline double weno_gather_face(const double* w, const double* Rk,
                                                   const double* G) {
  constexpr double c1 = RIGHT ? 1.0 : 5.0, c3 = RIGHT ? 5.0 : 1.0;
  constexpr int E = 4 - F;
  const double* y = w + F;
  const double u1 = weno_sub_coef<RIGHT>(1, 0) * y[0] + weno_sub_coef<RIGHT>(1, 1) * y[1] +
                    weno_sub_coef<RIGHT>(1, 2) * y[2];
  const double u2 = weno_sub_coef<RIGHT>(2, 1) * y[1] + weno_sub_coef<RIGHT>(2, 2) * y[2] +
                    weno_sub_coef<RIGHT>(2, 3) * y[3];
  const double u3 = weno_sub_coef<RIGHT>(3, 2) * y[2] + weno_sub_coef<RIGHT>(3, 3) * y[3] +
                    weno_sub_coef<RIGHT>(3, 4) * y[4];
  const double W1 = c1 * Rk[F], W2 = 10.0 * Rk[F + 1], W3 = c3 * Rk[F + 2];
  const double rD = fast_rcp(W1 + W2 + W3);
  const double sv = (W1 * u1 + W2 * u2 + W3 * u3) * rD;
  double num = 0.0;
  if constexpr (F >= 2 && F <= 4) num += W1 * G[F] * (u1 - sv);
  if constexpr (F + 1 >= 2 && F + 1 <= 4) num += W2 * G[F + 1] * (u2 - sv);
  if constexpr (F + 2 >= 2 && F + 2 <= 4) num += W3 * G[F + 2] * (u3 - sv);
  if constexpr (weno_sub_coef<RIGHT>(1, E) != 0.0) num += W1 * weno_sub_coef<RIGHT>(1, E);
  if constexpr (weno_sub_coef<RIGHT>(2, E) != 0.0) num += W2 * weno_sub_coef<RIGHT>(2, E);
  if constexpr (weno_sub_coef<RIGHT>(3, E) != 0.0) num += W3 * weno_sub_coef<RIGHT>(3, E);
  return num * rD;
}

template <bool RIGHT>
__host__ __device__ inline double weno_gather(const double* w, const double* m, double cdq) {
  double B[7], Rk[7], G[5];
#pragma unroll
  for (int c = 0; c < 7; ++c) {
    B[c] = weno_beta<double>(w[c], w[c + 1], w[c + 2]);
    Rk[c] = fast_rcp(16.0 * B[c] * B[c]);
  }
  G[2] = -2.0 * beta_dc(w[2], w[3], w[4]) * fast_rcp(B[2]);
  G[3] = -2.0 * beta_db(w[3], w[4], w[5]) * fast_rcp(B[3]);
  G[4] = -2.0 * beta_da(w[4], w[5], w[6]) * fast_rcp(B[4]);
  G[0] = G[1] = 0.0;  // never read
  // face f is the lower face of cell j + f - 1 (C > 0) or j + f - 2
  constexpr int o0 = RIGHT ? -1 : -2;
  double acc = cdq * (m[o0 + 3] - m[o0 + 2]) * weno_gather_face<RIGHT, 0>(w, Rk, G);
  acc += cdq * (m[o0 + 4] - m[o0 + 3]) * weno_gather_face<RIGHT, 1>(w, Rk, G);
  acc += cdq * (m[o0 + 5] - m[o0 + 4]) * weno_gather_face<RIGHT, 2>(w, Rk, G);
  acc += cdq * (m[o0 + 6] - m[o0 + 5]) * weno_gather_face<RIGHT, 3>(w, Rk, G);
  acc += cdq * (m[o0 + 7] - m[o0 + 6]) * weno_gather_face<RIGHT, 4>(w, Rk, G);
  return acc;
}

// Gather of the advection adjoint along one axis for the cell at window
// centre: w[0..8] = u at offsets -4..4, m[0..6] = mu at offsets -3..3,
// cdq = C/dq.  Upwind (C > 0): the lower face of cell k is the right face of
// cell k-1, stencil k-3..k+1, for k = j-1..j+3; otherwise the left face of
// cell k, stencil k-2..k+2, for k = j-2..j+2.
template <int R>
__host__ __device__ inline double axis_adjoint(const double* w, const double* m, double cdq) {
  if constexpr (R == RC_WENO5) {
    if (cdq > 0.0) return weno_gather<true>(w, m, cdq);
    if (cdq < 0.0) return weno_gather<false>(w, m, cdq);
    return 0.0;
  }
  double acc = 0.0;
  if (cdq > 0.0) {
#pragma unroll
    for (int o = -1; o <= 3; ++o) {
      const double sk = cdq * (m[o + 3] - m[o + 2]);
      acc += sk * face_partial<R>(w, o + 1, true);
    }
  } else if (cdq < 0.0) {
#pragma unroll
    for (int o = -2; o <= 2; ++o) {
      const double sk = cdq * (m[o + 3] - m[o + 2]);
      acc += sk * face_partial<R>(w, o + 2, false);
    }
  }
  return acc;
}

struct AdjParams {
  int n1, n2, P, n_p1, n_p2, n_p3, layout, p_dim, source;
  double inv_dq1, inv_dq2, inv_tau, measure, mass, kB;
  double p1s, dp1, p2s, dp2, p3s, dp3;
  const double* u;
  const double* mom;  // moment block of u
  const double* lam;
  const double* base;
  const double* base2;
  double* out;
  double a, a2, b, c;
  double* gq;        // [5][Q] moment weights of the collision Jacobian
  double* tau_grad;  // [Q] or null
};

__host__ __device__ inline long adj_index(const AdjParams& ap, int i1, int i2, int p) {
  return ap.layout ? ((long)i2 * ap.P + p) * ap.n1 + i1 : ((long)i2 * ap.n1 + i1) * ap.P + p;
}

// Per q-cell: reductions over p and the chain rule to the raw moments.
// The sums R0 = sum mu f_MB, R1k = sum mu f_MB (v_k - u_k), R2 = sum mu f_MB
// |v - u|^2 and S = sum mu (f_MB - f) over p = p0, p0 + step, ... < P.
__host__ __device__ inline void adj_reduce_sums(const AdjParams& ap, int q, int p0, int step,
                                                double r[6]) {
  const int Q = ap.n1 * ap.n2;
  const int i1 = q % ap.n1, i2 = q / ap.n1;
  const double* b = ap.mom + MOM_BGK * Q + q;
  const double A = b[0], u1 = b[Q], u2 = b[2 * Q], u3 = b[3 * Q];
  const double beta = b[6 * Q];
  const int d = ap.p_dim;
  double R0 = 0, R11 = 0, R12 = 0, R13 = 0, R2 = 0, S = 0;
  int j1 = p0 % ap.n_p1, j2 = (p0 / ap.n_p1) % ap.n_p2, j3 = p0 / (ap.n_p1 * ap.n_p2);
  for (int p = p0; p < ap.P; p += step) {
    const double v1 = centre(ap.p1s, j1, ap.dp1), v2 = centre(ap.p2s, j2, ap.dp2);
    const double v3 = centre(ap.p3s, j3, ap.dp3);
    const double d1 = v1 - u1, d2 = v2 - u2, d3 = v3 - u3;
    double q2 = d1 * d1;
    if (d >= 2) q2 += d2 * d2;
    if (d >= 3) q2 += d3 * d3;
    const double fmb = A * exp(-beta * q2);
    const long idx = adj_index(ap, i1, i2, p);
    const double mu = ap.lam[idx];
    const double w = mu * fmb;
    R0 += w;
    R11 += w * d1;
    R12 += w * d2;
    R13 += w * d3;
    R2 += w * q2;
    S += mu * (fmb - ap.u[idx]);
    j1 += step;
    while (j1 >= ap.n_p1) {
      j1 -= ap.n_p1;
      if (++j2 == ap.n_p2) {
        j2 = 0;
        ++j3;
      }
    }
  }
  r[0] = R0;
  r[1] = R11;
  r[2] = R12;
  r[3] = R13;
  r[4] = R2;
  r[5] = S;
}

// The chain rule from the sums of q-cell q to its moment weights (and the
// tau gradient).
__host__ __device__ inline void adj_reduce_finish(const AdjParams& ap, int q, const double r[6]) {
  const int Q = ap.n1 * ap.n2;
  const double* b = ap.mom + MOM_BGK * Q + q;
  const double u1 = b[Q], u2 = b[2 * Q], u3 = b[3 * Q], T = b[4 * Q], n = b[5 * Q];
  const int d = ap.p_dim;
  const double R0 = r[0], R11 = r[1], R12 = r[2], R13 = r[3], R2 = r[4], S = r[5];
  const double eps = 1e-30, m = ap.mass, kB = ap.kB, ne = n + eps;
  // derivatives of G = sum_p mu f_MB with respect to n, u_k, T
  const double Gn = R0 / n;
  const double mkT = m / (kB * T);
  const double Gu1 = mkT * R11;
  const double Gu2 = d >= 2 ? mkT * R12 : 0.0;
  const double Gu3 = d >= 3 ? mkT * R13 : 0.0;
  const double GT = -0.5 * d / T * R0 + 0.5 * mkT / T * R2;
  // T(E, n, u): dT/dE, dT/du_k, dT/dn at fixed E and u
  const double TE = 2.0 * m / (d * ne);
  const double usq = u1 * u1 + u2 * u2 + u3 * u3;
  const double Tn = -m * usq / (d * ne) - (T - eps) / ne;
  const double Tu = -2.0 * n * m / (d * ne);
  const double H1 = Gu1 + GT * Tu * u1, H2 = Gu2 + GT * Tu * u2, H3 = Gu3 + GT * Tu * u3;
  double* g = ap.gq + q;
  g[0] = Gn + GT * Tn - (H1 * u1 + H2 * u2 + H3 * u3) / ne;
  g[Q] = H1 / ne;
  g[2 * Q] = H2 / ne;
  g[3 * Q] = H3 / ne;
  g[4 * Q] = GT * TE;
  // dL/dtau = -(f_MB - f)/tau^2
  if (ap.tau_grad) ap.tau_grad[q] = -ap.c * ap.inv_tau * ap.inv_tau * S;
}

// The same for one q-cell, summed in plain p order.
__host__ __device__ inline void adj_reduce_cell(const AdjParams& ap, int q) {
  double r[6];
  adj_reduce_sums(ap, q, 0, 1, r);
  adj_reduce_finish(ap, q, r);
}

constexpr int ADJ_RW = 8;   // waves of a q1-line reduction workgroup
constexpr int ADJ_PCH = 8;  // p-chunks of the q1-line reduction

// The q1-line reduction splits p into ADJ_RW * ADJ_PCH interleaved streams:
// stream s = c * ADJ_RW + w takes p = s, s + ADJ_RW * ADJ_PCH, ...  Chunk c
// adds its streams in w order; the chunks are then added in c order.
__host__ __device__ inline void adj_reduce_chunk(const AdjParams& ap, int q, int c, double r[6]) {
  for (int k = 0; k < 6; ++k) r[k] = 0.0;
  for (int w = 0; w < ADJ_RW; ++w) {
    double x[6];
    adj_reduce_sums(ap, q, c * ADJ_RW + w, ADJ_RW * ADJ_PCH, x);
    for (int k = 0; k < 6; ++k) r[k] += x[k];
  }
}

#ifndef BX_NO_KERNELS  // host-only builds of the per-cell code leave the kernels out
// q1-line layout: workgroup (64 consecutive q1 cells of one q2 row, p-chunk
// c = blockIdx.y): lane = q1 cell (coalesced along the q1 lines), wave w =
// stream c * ADJ_RW + w; the waves' sums are added in wave order into
// part[c][6][Q].
__global__ __launch_bounds__(64 * ADJ_RW) void adj_bgk_reduce_lines_kernel(AdjParams ap,
                                                                          double* part) {
  __shared__ double red[6][ADJ_RW][64];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int nblk = (ap.n1 + 63) / 64;
  const int i1 = (blockIdx.x % nblk) * 64 + lane, i2 = blockIdx.x / nblk;
  const int c = blockIdx.y;
  const bool live = i1 < ap.n1;
  const int Q = ap.n1 * ap.n2;
  const int q = i2 * ap.n1 + (live ? i1 : 0);
  double r[6] = {0, 0, 0, 0, 0, 0};
  if (live) adj_reduce_sums(ap, q, c * ADJ_RW + wave, ADJ_RW * ADJ_PCH, r);
  for (int k = 0; k < 6; ++k) red[k][wave][lane] = r[k];
  __syncthreads();
  if (wave == 0 && live) {
    for (int k = 0; k < 6; ++k) {
      double x = 0.0;
      for (int w = 0; w < ADJ_RW; ++w) x += red[k][w][lane];
      part[((long)c * 6 + k) * Q + q] = x;
    }
  }
}

// The chunks' sums of every q-cell added in chunk order, then the chain rule.
__global__ __launch_bounds__(256) void adj_reduce_finish_kernel(AdjParams ap, const double* part) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int Q = ap.n1 * ap.n2;
  if (q >= Q) return;
  double r[6] = {0, 0, 0, 0, 0, 0};
  for (int c = 0; c < ADJ_PCH; ++c)
    for (int k = 0; k < 6; ++k) r[k] += part[((long)c * 6 + k) * Q + q];
  adj_reduce_finish(ap, q, r);
}

// Reference layout: a workgroup per q-cell, threads over p (contiguous),
// a butterfly within each wave and the waves added in order.
__global__ __launch_bounds__(256) void adj_bgk_reduce_kernel(AdjParams ap) {
  __shared__ double red[6][4];
  const int q = blockIdx.x;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double r[6];
  adj_reduce_sums(ap, q, threadIdx.x, 256, r);
  for (int k = 0; k < 6; ++k) {
    double x = r[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if (lane == 0) red[k][wave] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < 6; ++k) r[k] = ((red[k][0] + red[k][1]) + red[k][2]) + red[k][3];
    adj_reduce_finish(ap, q, r);
  }
}
#endif

// Cell (i1, i2, p) at memory offset t: out = a*base + b*mu + c*(J^T mu).
template <int R>
__host__ __device__ inline void adj_stage_at(const AdjParams& ap, int i1, int i2, int p, long t) {
  const long Qn = (long)ap.n1 * ap.n2;
  int j1, j2;
  const int j23 = fdivmod(p, ap.n_p1, 1.0f / float(ap.n_p1), j1);
  const int j3 = fdivmod(j23, ap.n_p2, 1.0f / float(ap.n_p2), j2);
  const double v1 = centre(ap.p1s, j1, ap.dp1), v2 = centre(ap.p2s, j2, ap.dp2);
  const double v3 = centre(ap.p3s, j3, ap.dp3);
  const double mu = ap.lam[t];
  double acc = 0.0;
  double w[9], m[7];
  // neighbours along q1 are s1 elements apart in memory, along q2 n1*P (both
  // layouts keep i2 outermost); offsets reach +-4 cells, one wrap for n >= 5
  const long s1 = ap.layout ? 1 : ap.P, s2 = (long)ap.n1 * ap.P;
  if (ap.n1 > 1) {
#pragma unroll
    for (int r = 0; r < 9; ++r) {
      const int i = ap.n1 >= 5 ? wrap_once(i1 + r - 4, ap.n1) : qm_mod(i1 + r - 4, ap.n1);
      w[r] = ap.u[t + (long)(i - i1) * s1];
      if (r >= 1 && r <= 7) m[r - 1] = ap.lam[t + (long)(i - i1) * s1];
    }
    acc += axis_adjoint<R>(w, m, v1 * ap.inv_dq1);
  }
  if (ap.n2 > 1) {
#pragma unroll
    for (int r = 0; r < 9; ++r) {
      const int i = ap.n2 >= 5 ? wrap_once(i2 + r - 4, ap.n2) : qm_mod(i2 + r - 4, ap.n2);
      w[r] = ap.u[t + (long)(i - i2) * s2];
      if (r >= 1 && r <= 7) m[r - 1] = ap.lam[t + (long)(i - i2) * s2];
    }
    acc += axis_adjoint<R>(w, m, v2 * ap.inv_dq2);
  }
  if (ap.source) {
    const int q = i2 * ap.n1 + i1;
    const double* g = ap.gq + q;
    const double e = 0.5 * ((v1 * v1 + v2 * v2) + v3 * v3);
    acc += ap.inv_tau * (ap.measure * (g[0] + g[Qn] * v1 + g[2 * Qn] * v2 + g[3 * Qn] * v3 +
                                       g[4 * Qn] * e) - mu);
  }
  double x = ap.b * mu + ap.c * acc;
  if (ap.base) x += ap.a * ap.base[t];
  if (ap.base2) x += ap.a2 * ap.base2[t];
  ap.out[t] = x;
}

// The same for cell t of memory order (either layout keeps i2 outermost).
template <int R>
__host__ __device__ inline void adj_stage_cell(const AdjParams& ap, long t) {
  const int slab = ap.n1 * ap.P;
  const int i2 = (int)(t / slab), x = (int)(t % slab);
  if (ap.layout) adj_stage_at<R>(ap, x % ap.n1, i2, x / ap.n1, t);
  else adj_stage_at<R>(ap, x / ap.P, i2, x % ap.P, t);
}

#ifndef BX_NO_KERNELS
// Grid: x over the n1*P cells of one q2 row (< 2^31, checked by the host), y = i2.
template <int R>
__global__ __launch_bounds__(256) void adj_stage_kernel(AdjParams ap) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int slab = ap.n1 * ap.P;
  if (x >= slab) return;
  const int i2 = blockIdx.y;
  const long t = (long)i2 * slab + x;
  if (ap.layout) adj_stage_at<R>(ap, x % ap.n1, i2, x / ap.n1, t);
  else adj_stage_at<R>(ap, x / ap.P, i2, x % ap.P, t);
}
#endif

// Parameters of one adjoint stage from the C-ABI structs.
__host__ inline AdjParams adj_params(const boltx_grid* g, const boltx_physics* ph,
                                     const boltx_adjoint* adj) {
  AdjParams ap;
  ap.n1 = g->n_q1;
  ap.n2 = g->n_q2;
  ap.n_p1 = g->n_p1;
  ap.n_p2 = g->n_p2;
  ap.n_p3 = g->n_p3;
  ap.P = g->n_p1 * g->n_p2 * g->n_p3;
  ap.layout = g->layout;
  ap.p_dim = ph->p_dim;
  ap.source = ph->source_enabled;
  ap.inv_dq1 = 1.0 / g->dq1;
  ap.inv_dq2 = 1.0 / g->dq2;
  ap.inv_tau = ph->source_enabled ? 1.0 / ph->tau : 0.0;
  ap.measure = g->dp1 * g->dp2 * g->dp3;
  ap.mass = ph->mass;
  ap.kB = ph->boltzmann_constant;
  ap.p1s = g->p1_start;
  ap.dp1 = g->dp1;
  ap.p2s = g->p2_start;
  ap.dp2 = g->dp2;
  ap.p3s = g->p3_start;
  ap.dp3 = g->dp3;
  ap.u = (const double*)adj->u;
  ap.mom = adj->mom_u;
  ap.lam = (const double*)adj->lam;
  ap.base = (const double*)adj->base;
  ap.base2 = (const double*)adj->base2;
  ap.out = (double*)adj->out;
  ap.a = adj->a;
  ap.a2 = adj->a2;
  ap.b = adj->b;
  ap.c = adj->c;
  ap.gq = (double*)adj->workspace;
  ap.tau_grad = adj->tau_grad;
  return ap;
}

}  // namespace bx
