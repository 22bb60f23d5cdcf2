// One Runge-Kutta stage of the nonlinear FVM step for a single phase-space
// cell, shared by the gfx950 kernels and the CPU harness of the tests.
//
// Reference: df_dt_fvm (bolt/lib/nonlinear/finite_volume/df_dt_fvm.py:33-339)
// evaluated inside one stage of timestep_fvm (fvm_operator.py:22-33):
//   L = (-(F1(i+1/2)-F1(i-1/2))/dq1 - (F2(j+1/2)-F2(j-1/2))/dq2)   (:88-89)
//     + BGK(u)                                                    (:91-97)
//     - ((dF_p1 + dF_p2) + dF_p3)                                 (:319-323)
//   out = a*un + b*(w + cdt*L)
#pragma once
#include "fvm_math.h"

namespace bx {

// A term supplied by the caller instead of the built-in one (a user
// callback evaluated on the device): the value at p-cell p of interior
// q-cell (i1, i2) is ptr[p*sp + i1*s1 + i2*s2] (element strides; a stride of 0
// broadcasts along that index).  Face arrays are read at i1 = n_q1 (i2 =
// n_q2) as well: the left face of the cell after the last interior one.
struct Coeff {
  const double* ptr;  // NULL: the built-in term
  long sp, s1, s2;
};

__host__ __device__ inline double coeff_at(const Coeff& c, long p, long i1, long i2) {
  return c.ptr[p * c.sp + i1 * c.s1 + i2 * c.s2];
}

struct StageParams {
  int n_q1, n_q2, n_p1, n_p2, n_p3, P, Q;
  double p1s, p2s, p3s, dp1, dp2, dp3;
  double inv_dq1, inv_dq2, inv_dp1, inv_dp2, inv_dp3;
  int riemann_q, riemann_p, source, p_dim;
  double mass, charge, kB, inv_tau;
  double a, b, cdt;
  double measure;  // dp3*dp2*dp1
  const void* u;
  const void* w;
  const void* un;
  void* out;
  const double* mom_in;
  double* mom_out;
  const double* fp1;  // [6][Q] fields seen by C_p1
  const double* fp2;
  const double* fp3;
  // non-periodic q axes: ghost q-rows follow the Q interior rows of every
  // f-sized array (left, right, bottom, top blocks of ng rows per q-line)
  int ng, gh1, gh2;
  // caller-supplied terms (generic kernel only): C_q1 at left q1-faces, C_q2
  // at bottom q2-faces, C_p1..3 at the left / bottom / back p-faces, an
  // additive source replacing BGK, the BGK collision time per cell
  Coeff cq1, cq2, cp1, cp2, cp3, src, tau;
  // [3][Q] currents of u (charge * mom_v1(f_left), mom_v2(f_bot), mom_v3(f)),
  // accumulated by the slab kernels, or NULL
  double* Jout;
};

// Per-q-cell state the stage needs besides f: the stencil neighbours of the
// q-cell (as q-cell indices) and the Maxwellian of its moments.
template <typename T>
struct QCell {
  int q;
  int nb1[7];  // rows of q-cells i1-3 .. i1+3 (same i2): periodic images or ghost rows
  int nb2[7];  // rows of q-cells i2-3 .. i2+3 (same i1)
  double pref; // n (m / 2 pi k T)^(p_dim/2)
  const T* e1; // exp(-m (v1-u1)^2 / 2kT) per p1 index
  const T* e2;
  const T* e3;
  double E[3], B[3][3];  // fields for C_p1 (row 0), C_p2 (row 1), C_p3 (row 2)
};

// Row of q-cell (j1, i2) where j1 may lie outside [0, n_q1): the periodic
// image, or the ghost row of a non-periodic axis (a stencil never reaches
// past the ng ghost cells; beyond them the index is clamped).
__host__ __device__ inline int q_row_1(const StageParams& sp, int j1, int i2) {
  if (j1 >= 0 && j1 < sp.n_q1) return j1 + sp.n_q1 * i2;
  if (!sp.gh1) return ((j1 % sp.n_q1 + sp.n_q1) % sp.n_q1) + sp.n_q1 * i2;
  const bool right = j1 >= sp.n_q1;
  int c = right ? j1 - sp.n_q1 : j1 + sp.ng;
  c = c < 0 ? 0 : (c >= sp.ng ? sp.ng - 1 : c);
  return sp.Q + (right ? sp.ng * sp.n_q2 : 0) + c * sp.n_q2 + i2;
}

// The same along q2 for q-cell (i1, j2).
__host__ __device__ inline int q_row_2(const StageParams& sp, int i1, int j2) {
  if (j2 >= 0 && j2 < sp.n_q2) return i1 + sp.n_q1 * j2;
  if (!sp.gh2) return i1 + sp.n_q1 * ((j2 % sp.n_q2 + sp.n_q2) % sp.n_q2);
  const bool top = j2 >= sp.n_q2;
  int c = top ? j2 - sp.n_q2 : j2 + sp.ng;
  c = c < 0 ? 0 : (c >= sp.ng ? sp.ng - 1 : c);
  return sp.Q + (sp.gh1 ? 2 * sp.ng * sp.n_q2 : 0) + (top ? sp.ng * sp.n_q1 : 0) +
         c * sp.n_q1 + i1;
}

template <typename T>
__host__ __device__ inline void qcell_neighbours(const StageParams& sp, int q, QCell<T>& c) {
  const int i1 = q % sp.n_q1, i2 = q / sp.n_q1;
  c.q = q;
  for (int k = 0; k < 7; ++k) {
    c.nb1[k] = q_row_1(sp, i1 + k - 3, i2);
    c.nb2[k] = q_row_2(sp, i1, i2 + k - 3);
  }
}

// Fields at cell q for the three C_p evaluations, advection_terms.py:179-186.
template <typename T>
__host__ __device__ inline void qcell_fields(const StageParams& sp, QCell<T>& c) {
  const double* f[3] = {sp.fp1, sp.fp2, sp.fp3};
  for (int k = 0; k < 3; ++k) {
    c.E[k] = f[k][k * sp.Q + c.q];  // C_p1 needs E1, C_p2 E2, C_p3 E3
    for (int j = 0; j < 3; ++j) c.B[k][j] = f[k][(3 + j) * sp.Q + c.q];
  }
}

// Raw moments (n, n u, E) of q-cell q -> Maxwellian prefactor, and the
// separable factors written to e1/e2/e3 by the caller from bgk_cell().
__host__ __device__ inline BgkCell qcell_bgk(const StageParams& sp, int q) {
  const double* m = sp.mom_in;
  return bgk_cell(m[q], m[sp.Q + q], m[2 * sp.Q + q], m[3 * sp.Q + q], m[4 * sp.Q + q],
                  sp.mass, sp.p_dim);
}

// Flux difference along one p-axis from a seven-point window of the q-cell's
// block (periodic in p), df_dt_fvm.py:274-321.
template <int RP, typename T>
__host__ __device__ inline T p_flux_diff(const T* blk, int p, int j, int n, int stride, T C,
                                         int riemann) {
  T a[7];
  for (int k = 0; k < 7; ++k) {
    const int jj = ((j + k - 3) % n + n) % n;
    a[k] = blk[p + (jj - j) * stride];
  }
  return flux_diff<RP, T>(a, C, riemann, (T*)nullptr);
}

// The same with a caller-supplied C_p along the axis: C at the cell's own
// lower face and at the lower face of the next cell along the axis (the
// periodic shift of df_dt_fvm.py:306-308 wraps the last cell's upper face to
// the first cell's lower face).
template <int RP, typename T>
__host__ __device__ inline T p_flux_diff_coeff(const T* blk, int p, int j, int n, int stride,
                                               const Coeff& C, int i1, int i2, int riemann) {
  T a[7];
  for (int k = 0; k < 7; ++k) {
    const int jj = ((j + k - 3) % n + n) % n;
    a[k] = blk[p + (jj - j) * stride];
  }
  const int jr = j + 1 == n ? 0 : j + 1;
  const T CL = T(coeff_at(C, p, i1, i2));
  const T CR = T(coeff_at(C, p + (jr - j) * stride, i1, i2));
  return flux_diff_var<RP, T>(a, CL, CR, riemann);
}

// The stage value of phase-space cell (q, p).  ``blk`` is the q-cell's P
// values of u (needed only when p_advection is on).
template <typename T, int RQ, int RP, bool PADV>
__host__ __device__ inline T stage_cell(const StageParams& sp, const QCell<T>& c, int p,
                                        const T* blk, double v[3]) {
  const T* u = (const T*)sp.u;
  const int j1 = p % sp.n_p1;
  const int j23 = p / sp.n_p1;
  const int j2 = j23 % sp.n_p2;
  const int j3 = j23 / sp.n_p2;
  v[0] = centre(sp.p1s, j1, sp.dp1);
  v[1] = centre(sp.p2s, j2, sp.dp2);
  v[2] = centre(sp.p3s, j3, sp.dp3);
  const long P = sp.P;
  const T f = u[(long)c.q * P + p];
  const int i1 = c.q % sp.n_q1, i2 = c.q / sp.n_q1;
  constexpr int r = recon_reach(RQ) + 1;  // neighbours read on each side
  T d = T(0);
  // q1 and q2 fluxes; built-in C_q1 = v1, C_q2 = v2 (advection_terms.py:64-97),
  // else the caller's face values
  T t1 = T(0), t2 = T(0);
  if (sp.n_q1 > 1) {
    T a[7];
    for (int k = 0; k < 7; ++k) a[k] = T(0);
    for (int k = 3 - r; k <= 3 + r; ++k) a[k] = k == 3 ? f : u[(long)c.nb1[k] * P + p];
    if (sp.cq1.ptr)
      t1 = flux_diff_var<RQ, T>(a, T(coeff_at(sp.cq1, p, i1, i2)),
                                T(coeff_at(sp.cq1, p, i1 + 1, i2)), sp.riemann_q);
    else
      t1 = flux_diff<RQ, T>(a, T(v[0]), sp.riemann_q, (T*)nullptr);
  }
  if (sp.n_q2 > 1) {
    T a[7];
    for (int k = 0; k < 7; ++k) a[k] = T(0);
    for (int k = 3 - r; k <= 3 + r; ++k) a[k] = k == 3 ? f : u[(long)c.nb2[k] * P + p];
    if (sp.cq2.ptr)
      t2 = flux_diff_var<RQ, T>(a, T(coeff_at(sp.cq2, p, i1, i2)),
                                T(coeff_at(sp.cq2, p, i1, i2 + 1)), sp.riemann_q);
    else
      t2 = flux_diff<RQ, T>(a, T(v[1]), sp.riemann_q, (T*)nullptr);
  }
  d = -t1 * T(sp.inv_dq1) - t2 * T(sp.inv_dq2);
  if (sp.src.ptr) {
    d = d + T(coeff_at(sp.src, p, i1, i2));  // df_dt_fvm.py:91-97, the caller's source
  } else if (sp.source) {
    // f_MB = ((n c^(p_dim/2) * e1) * e2) * e3 (collision_operator.py:52-65)
    T fmb = T(c.pref) * c.e1[j1];
    if (sp.p_dim >= 2) fmb = fmb * c.e2[j2];
    if (sp.p_dim >= 3) fmb = fmb * c.e3[j3];
    if (sp.tau.ptr)
      d = d + (fmb - f) / T(coeff_at(sp.tau, p, i1, i2));  // collision_operator.py:119-128
    else
      d = d + (fmb - f) * T(sp.inv_tau);
  }
  if constexpr (PADV) if (sp.cp1.ptr) {
    // the caller's C_p1..3 at the p-faces (df_dt_fvm.py:251-321)
    T dp = T(0);
    if (sp.n_p1 > 1)
      dp = p_flux_diff_coeff<RP, T>(blk, p, j1, sp.n_p1, 1, sp.cp1, i1, i2, sp.riemann_p) *
           T(sp.inv_dp1);
    if (sp.n_p2 > 1)
      dp = dp + p_flux_diff_coeff<RP, T>(blk, p, j2, sp.n_p2, sp.n_p1, sp.cp2, i1, i2,
                                         sp.riemann_p) * T(sp.inv_dp2);
    if (sp.n_p3 > 1)
      dp = dp + p_flux_diff_coeff<RP, T>(blk, p, j3, sp.n_p3, sp.n_p1 * sp.n_p2, sp.cp3, i1,
                                         i2, sp.riemann_p) * T(sp.inv_dp3);
    d = d - dp;
  } else {
    // C_p at the p-faces, cell-centred fields (advection_terms.py:184-186)
    const double em = sp.charge / sp.mass;
    T dp = T(0);
    if (sp.n_p1 > 1) {
      const double C = em * (c.E[0] + v[1] * c.B[0][2] - v[2] * c.B[0][1]);
      dp = p_flux_diff<RP, T>(blk, p, j1, sp.n_p1, 1, T(C), sp.riemann_p) * T(sp.inv_dp1);
    }
    if (sp.n_p2 > 1) {
      const double C = em * (c.E[1] + v[2] * c.B[1][0] - v[0] * c.B[1][2]);
      dp = dp + p_flux_diff<RP, T>(blk, p, j2, sp.n_p2, sp.n_p1, T(C), sp.riemann_p) *
                    T(sp.inv_dp2);
    }
    if (sp.n_p3 > 1) {
      const double C = em * (c.E[2] + v[0] * c.B[2][1] - v[1] * c.B[2][0]);
      dp = dp + p_flux_diff<RP, T>(blk, p, j3, sp.n_p3, sp.n_p1 * sp.n_p2, T(C),
                                   sp.riemann_p) * T(sp.inv_dp3);
    }
    d = d - dp;
  }
  const long idx = (long)c.q * P + p;
  const T wv = ((const T*)sp.w)[idx];
  const T t = wv + d * T(sp.cdt);
  if (sp.a != 0.0) return T(sp.a) * ((const T*)sp.un)[idx] + T(sp.b) * t;
  return sp.b == 1.0 ? t : T(sp.b) * t;
}

}  // namespace bx
