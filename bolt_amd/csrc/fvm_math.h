// Per-cell arithmetic of the nonlinear FVM step (reconstruction, Riemann
// fluxes, BGK Maxwellian, RK combination).  Every function is
// __host__ __device__ so the same code runs in the gfx950 kernels and in the
// CPU harness the tests build from this header.
//
// Reference semantics (QuazarTech/Bolt):
//   minmod   bolt/lib/nonlinear/finite_volume/reconstruction/minmod.py:4-71
//   ppm      bolt/lib/nonlinear/finite_volume/reconstruction/ppm.py:4-100
//   weno5    bolt/lib/nonlinear/finite_volume/reconstruction/weno5.py:4-98
//   riemann  bolt/lib/nonlinear/finite_volume/riemann.py:46-49 (upwind select)
//   fluxes   bolt/lib/nonlinear/finite_volume/df_dt_fvm.py:56-97, 251-323
//   BGK      bolt/src/nonrelativistic_boltzmann/collision_operator.py:15-128
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bx {

enum Recon : int { RC_PC = 0, RC_MINMOD = 1, RC_PPM = 2, RC_WENO5 = 3 };
enum Riemann : int { RS_UPWIND = 0, RS_LLF = 1 };

// Cells a reconstruction reaches on each side of the cell it reconstructs.
__host__ __device__ constexpr int recon_reach(int r) {
  return r == RC_PC ? 0 : (r == RC_MINMOD ? 1 : 2);
}

template <typename T> __host__ __device__ inline T amin(T a, T b) { return a < b ? a : b; }
__host__ __device__ inline double aabs(double a) { return __builtin_fabs(a); }
__host__ __device__ inline float aabs(float a) { return __builtin_fabsf(a); }

// 1/x to within an ulp.  On the device: hardware reciprocal refined by Newton
// steps (two for double); on the host: plain division.
__host__ __device__ inline double fast_rcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
#else
  return 1.0 / x;
#endif
}
__host__ __device__ inline float fast_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  float r = __builtin_amdgcn_rcpf(x);
  float e = __builtin_fmaf(-x, r, 1.0f);
  return __builtin_fmaf(r, e, r);
#else
  return 1.0f / x;
#endif
}

// ---------------------------------------------------------------------------
// minmod.py:4-62: MC-limited slope minmod(2*bwd, 0.5*(bwd+fwd), 2*fwd) with
// af.sign(x) = 1 for x < 0; the 0.25*|sx+sy|*(sx+sz) factor is evaluated as
// written (it is exact: products of small integers).
template <typename T>
__host__ __device__ inline T mc_slope(T ym, T y, T yp) {
  const T fwd = yp - y;
  const T bwd = y - ym;
  const T cen = bwd + fwd;
  const T x = T(2) * bwd, c = T(0.5) * cen, z = T(2) * fwd;
  const T mn = amin(amin(aabs(x), aabs(c)), aabs(z));
  const T sx = x < T(0) ? T(-1) : T(1);
  const T sy = c < T(0) ? T(-1) : T(1);
  const T sz = z < T(0) ? T(-1) : T(1);
  return T(0.25) * aabs(sx + sy) * (sx + sz) * mn;
}

// ppm.py:58-80: the limited slope of the centre cell y of ym, y, yp.
template <typename T>
__host__ __device__ inline T ppm_dq(T ym, T y, T yp) {
  // no contraction: a*b - c*d forms would otherwise round differently
  // wherever the compiler fuses the other product
#pragma clang fp contract(off)
  const T dm = T(2) * (y - ym), dp = T(2) * (yp - y);
  const T D = T(0.5) * (yp - ym);
  return (dp * dm <= T(0)) ? T(0)
         : (D > T(0) ? T(1) : T(-1)) * amin(aabs(D), amin(aabs(dm), aabs(dp)));
}

// ppm.py:82-98: the two faces of the centre cell y2 of y1, y2, y3 from the
// limited slopes DQ1, DQ2, DQ3 of the three cells.
template <typename T>
__host__ __device__ inline void ppm_faces_dq(T y1, T y2, T y3, T DQ1, T DQ2, T DQ3, T& left,
                                             T& right) {
#pragma clang fp contract(off)
  const T sixth = T(1.0 / 6.0);
  T l = T(0.5) * (y2 + y1) - (DQ2 - DQ1) * sixth;
  T r = T(0.5) * (y3 + y2) - (DQ3 - DQ2) * sixth;
  const bool corr1 = (r - y2) * (y2 - l) <= T(0);
  const T qd = r - l;
  const T qe = T(6) * (y2 - T(0.5) * (r + l));
  const bool corr2 = qd * (qd - qe) < T(0);
  const bool corr3 = qd * (qd + qe) < T(0);
  if (corr1) { l = y2; r = y2; }
  if (corr2) l = T(3) * y2 - T(2) * r;
  else if (corr3) r = T(3) * y2 - T(2) * l;
  left = l;
  right = r;
}

// ppm.py:57-98 on the five-cell stencil y0..y4 of the centre cell y2: the
// arithmetic of ppm_faces_dq on ppm_dq of y1, y2, y3, written out in one
// piece, which the FP64 stage kernels schedule better (the two forms agree
// bit for bit).
template <typename T>
__host__ __device__ inline void ppm_faces(T y0, T y1, T y2, T y3, T y4, T& left, T& right) {
#pragma clang fp contract(off)
  const T d0 = T(2) * (y1 - y0), d1 = T(2) * (y2 - y1);
  const T d2 = T(2) * (y3 - y2), d3 = T(2) * (y4 - y3);
  const T D1 = T(0.5) * (y2 - y0), D2 = T(0.5) * (y3 - y1), D3 = T(0.5) * (y4 - y2);
  const T DQ1 = (d1 * d0 <= T(0)) ? T(0)
              : (D1 > T(0) ? T(1) : T(-1)) * amin(aabs(D1), amin(aabs(d0), aabs(d1)));
  const T DQ2 = (d2 * d1 <= T(0)) ? T(0)
              : (D2 > T(0) ? T(1) : T(-1)) * amin(aabs(D2), amin(aabs(d1), aabs(d2)));
  const T DQ3 = (d3 * d2 <= T(0)) ? T(0)
              : (D3 > T(0) ? T(1) : T(-1)) * amin(aabs(D3), amin(aabs(d2), aabs(d3)));
  const T sixth = T(1.0 / 6.0);
  T l = T(0.5) * (y2 + y1) - (DQ2 - DQ1) * sixth;
  T r = T(0.5) * (y3 + y2) - (DQ3 - DQ2) * sixth;
  const bool corr1 = (r - y2) * (y2 - l) <= T(0);
  const T qd = r - l;
  const T qe = T(6) * (y2 - T(0.5) * (r + l));
  const bool corr2 = qd * (qd - qe) < T(0);
  const bool corr3 = qd * (qd + qe) < T(0);
  if (corr1) { l = y2; r = y2; }
  if (corr2) l = T(3) * y2 - T(2) * r;
  else if (corr3) r = T(3) * y2 - T(2) * l;
  left = l;
  right = r;
}

// weno5.py:58-71: smoothness indicator in the expanded form the reference
// uses, evaluated in the reference's order without FMA contraction (its
// rounding noise sets the nonlinear weights on smooth data, so the FP64 path
// reproduces it term by term).
__host__ __device__ inline double weno_beta_literal(double a, double b, double c) {
#pragma clang fp contract(off)
  const double s = (4.0 / 3.0) * a * a - (19.0 / 3.0) * a * b + (25.0 / 3.0) * b * b
                 + (11.0 / 3.0) * a * c - (31.0 / 3.0) * b * c + (10.0 / 3.0) * c * c;
  return s + 1e-17 * (1.0 + aabs(a) + aabs(b) + aabs(c));
}

// The same indicator as a sum of squares, 13/12 (a-2b+c)^2 + 1/4 (a-4b+3c)^2
// (algebraically identical to the expanded form); used by the FP32 bulk path
// where the expanded form's cancellation would lose every digit.
__host__ __device__ inline float weno_beta_sos(float a, float b, float c) {
  const float s1 = a - 2.0f * b + c;
  const float s2 = a - 4.0f * b + 3.0f * c;
  return (13.0f / 12.0f) * s1 * s1 + 0.25f * s2 * s2
       + 1e-17f * (1.0f + aabs(a) + aabs(b) + aabs(c));
}

template <typename T> __host__ __device__ inline T weno_beta(T a, T b, T c);
template <> __host__ __device__ inline double weno_beta<double>(double a, double b, double c) {
  return weno_beta_literal(a, b, c);
}
template <> __host__ __device__ inline float weno_beta<float>(float a, float b, float c) {
  return weno_beta_sos(a, b, c);
}

// weno5.py:73-96 for one face of the centre cell y2 given its three
// indicators: right face (w = 1/16, 5/8, 5/16) or left face (5/16, 5/8, 1/16).
template <typename T>
__host__ __device__ inline T weno_combine(T y0, T y1, T y2, T y3, T y4,
                                          T b1, T b2, T b3, bool right) {
  const T r1 = fast_rcp(T(16) * b1 * b1);
  const T r2 = fast_rcp(T(8) * b2 * b2);
  const T r3 = fast_rcp(T(16) * b3 * b3);
  const T w1 = right ? r1 : T(5) * r1;
  const T w2 = T(5) * r2;
  const T w3 = right ? T(5) * r3 : r3;
  T u1, u2, u3;
  if (right) {
    u1 = T(0.375) * y0 - T(1.25) * y1 + T(1.875) * y2;
    u2 = T(-0.125) * y1 + T(0.75) * y2 + T(0.375) * y3;
    u3 = T(0.375) * y2 + T(0.75) * y3 - T(0.125) * y4;
  } else {
    u1 = T(-0.125) * y0 + T(0.75) * y1 + T(0.375) * y2;
    u2 = T(0.375) * y1 + T(0.75) * y2 - T(0.125) * y3;
    u3 = T(1.875) * y2 - T(1.25) * y3 + T(0.375) * y4;
  }
  const T num = w1 * u1 + w2 * u2 + w3 * u3;
  const T den = w1 + w2 + w3;
  return num * fast_rcp(den);
}

// Face value of the centre cell y2 of the window y0..y4: its right face
// (right = true) or its left face.
template <int R, typename T>
__host__ __device__ inline T face_value(T y0, T y1, T y2, T y3, T y4, bool right) {
  if constexpr (R == RC_PC) {
    return y2;
  } else if constexpr (R == RC_MINMOD) {
    const T s = mc_slope(y1, y2, y3);
    return right ? y2 + T(0.5) * s : y2 - T(0.5) * s;
  } else if constexpr (R == RC_PPM) {
    T l, r;
    ppm_faces(y0, y1, y2, y3, y4, l, r);
    return right ? r : l;
  } else {
    const T b1 = weno_beta<T>(y0, y1, y2);
    const T b2 = weno_beta<T>(y1, y2, y3);
    const T b3 = weno_beta<T>(y2, y3, y4);
    return weno_combine(y0, y1, y2, y3, y4, b1, b2, b3, right);
  }
}

// Flux difference F(i+1/2) - F(i-1/2) for cell i from the seven values
// a[k] = f[i-3+k] and a face velocity C that does not vary along the axis.
// Left face i-1/2: minus state = right face of cell i-1 (window a0..a4),
// plus state = left face of cell i (a1..a5); right face i+1/2: a1..a5 and
// a2..a6 (df_dt_fvm.py:58-89).  Upwind takes the minus state when C > 0 and
// the plus state otherwise (riemann.py:46-49), so only the upwind state of
// each face is reconstructed.  LLF evaluates 0.5*(C*(fL+fR) - |C|*(fR-fL))
// with both states.  If ``left_state`` is non-null the upwind state of the
// left face is stored there (for the current moments).
template <int R, typename T>
__host__ __device__ inline T flux_diff(const T a[7], T C, int riemann, T* left_state) {
  if (riemann == RS_LLF) {
    const T mL = face_value<R>(a[0], a[1], a[2], a[3], a[4], true);
    const T pL = face_value<R>(a[1], a[2], a[3], a[4], a[5], false);
    const T mR = face_value<R>(a[1], a[2], a[3], a[4], a[5], true);
    const T pR = face_value<R>(a[2], a[3], a[4], a[5], a[6], false);
    const T aC = aabs(C);
    const T FL = T(0.5) * (C * (mL + pL) - aC * (pL - mL));
    const T FR = T(0.5) * (C * (mR + pR) - aC * (pR - mR));
    if (left_state) *left_state = C > T(0) ? mL : pL;
    return FR - FL;
  }
  const bool pos = C > T(0);
  // upwind window of the left face starts at a0 (C > 0) or a1 (C <= 0)
  const T w0 = pos ? a[0] : a[1];
  const T w1 = pos ? a[1] : a[2];
  const T w2 = pos ? a[2] : a[3];
  const T w3 = pos ? a[3] : a[4];
  const T w4 = pos ? a[4] : a[5];
  const T w5 = pos ? a[5] : a[6];
  const T sL = face_value<R>(w0, w1, w2, w3, w4, pos);
  const T sR = face_value<R>(w1, w2, w3, w4, w5, pos);
  if (left_state) *left_state = sL;
  return C * sR - C * sL;
}

// Flux difference F(i+1/2) - F(i-1/2) as flux_diff, with face velocities
// that differ between the two faces (a caller-supplied C: CL at the left
// face i-1/2, CR at the right face i+1/2), df_dt_fvm.py:67-89 and
// riemann.py:46-49 with C evaluated per face.  Both states of each face are
// reconstructed.
template <int R, typename T>
__host__ __device__ inline T flux_diff_var(const T a[7], T CL, T CR, int riemann) {
  const T mL = face_value<R>(a[0], a[1], a[2], a[3], a[4], true);
  const T pL = face_value<R>(a[1], a[2], a[3], a[4], a[5], false);
  const T mR = face_value<R>(a[1], a[2], a[3], a[4], a[5], true);
  const T pR = face_value<R>(a[2], a[3], a[4], a[5], a[6], false);
  T FL, FR;
  if (riemann == RS_LLF) {
    FL = T(0.5) * (CL * (mL + pL) - aabs(CL) * (pL - mL));
    FR = T(0.5) * (CR * (mR + pR) - aabs(CR) * (pR - mR));
  } else {
    FL = CL * (CL > T(0) ? mL : pL);
    FR = CR * (CR > T(0) ? mR : pR);
  }
  return FR - FL;
}

// collision_operator.py:102-116: bulk velocity and temperature from the raw
// moments (n, n*u1, n*u2, n*u3, E) of one q-cell, with the 1e-30 floors.
struct BgkCell {
  double n, u1, u2, u3, T;
};

__host__ __device__ inline BgkCell bgk_cell(double n, double m1, double m2, double m3,
                                            double E, double mass, int p_dim) {
  const double eps = 1e-30;
  BgkCell c;
  c.n = n;
  c.u1 = m1 / (n + eps);
  c.u2 = m2 / (n + eps);
  c.u3 = m3 / (n + eps);
  const double nm = n * mass;
  c.T = (1.0 / p_dim) * (2.0 * (E * mass) - nm * (c.u1 * c.u1) - nm * (c.u2 * c.u2)
                         - nm * (c.u3 * c.u3)) / (n + eps) + eps;
  return c;
}

// collision_operator.py:49-65: prefactor n*(m/(2 pi k T))^(p_dim/2) and the
// separable factor exp(-m (v - u)^2 / (2 k T)) along one velocity axis.
__host__ __device__ inline double maxwell_prefactor(const BgkCell& c, double mass, double kB,
                                                    int p_dim) {
  const double r = mass / (2.0 * 3.141592653589793 * kB * c.T);
  if (p_dim == 3) return c.n * pow(r, 1.5);
  if (p_dim == 2) return c.n * r;
  return c.n * sqrt(r);
}

__host__ __device__ inline double maxwell_factor(double v, double u, double T, double mass,
                                                 double kB) {
  const double d = v - u;
  return exp(-mass * (d * d) / (2.0 * kB * T));
}

// A moment block holds, per q-cell, the five built-in moments (density,
// mom_v1_bulk, mom_v2_bulk, mom_v3_bulk, energy; rows 0-4) followed by the
// BGK parameters derived from them (rows 5-11): prefactor n (m/2 pi k T)^(p_dim/2),
// u1, u2, u3, T, n, m/(2 k T).
constexpr int MOM_ROWS = 12;
constexpr int MOM_BGK = 5;

__host__ __device__ inline void write_moment_block(double* mom, int Q, int qc, const double m[5],
                                                   double mass, double kB, int p_dim) {
  for (int k = 0; k < 5; ++k) mom[k * Q + qc] = m[k];
  const BgkCell c = bgk_cell(m[0], m[1], m[2], m[3], m[4], mass, p_dim);
  double* b = mom + MOM_BGK * Q + qc;
  b[0] = maxwell_prefactor(c, mass, kB, p_dim);
  b[Q] = c.u1;
  b[2 * Q] = c.u2;
  b[3 * Q] = c.u3;
  b[4 * Q] = c.T;
  b[5 * Q] = c.n;
  b[6 * Q] = mass / (2.0 * kB * c.T);
}

// Sum of the five moment partials of one q-cell over nchunk chunks,
// part[(c * 5 + k) * stride], in chunk order for each k (the same result as
// the plain loop).
__host__ __device__ inline void sum_moment_partials(const double* part, int nchunk, long stride,
                                                    double m[5]) {
  double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  int c = 0;
  // eight chunks' loads in flight at a time; the sums still run in chunk order
  for (; c + 8 <= nchunk; c += 8) {
    double v[8][5];
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int k = 0; k < 5; ++k) v[j][k] = part[((long)(c + j) * 5 + k) * stride];
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int k = 0; k < 5; ++k) s[k] += v[j][k];
  }
  for (; c < nchunk; ++c)
#pragma unroll
    for (int k = 0; k < 5; ++k) s[k] += part[((long)c * 5 + k) * stride];
#pragma unroll
  for (int k = 0; k < 5; ++k) m[k] = s[k];
}

// Periodic index: i mod n in [0, n).
__host__ __device__ inline int qm_mod(int i, int n) {
  i %= n;
  return i < 0 ? i + n : i;
}

// The same for -n <= i < 2n (one conditional wrap).
__host__ __device__ inline int wrap_once(int i, int n) {
  return i < 0 ? i + n : (i >= n ? i - n : i);
}

// x / d (returned) and x % d (in rem) for 0 <= x < 2^24, 0 < d < 2^24, through
// the FP32 reciprocal rcp_d = 1/d: the quotient estimate is off by at most
// one and is corrected, instead of an integer division sequence.
__host__ __device__ inline int fdivmod(int x, int d, float rcp_d, int& rem) {
  int qt = int(float(x) * rcp_d);
  int r = x - qt * d;
  if (r < 0) {
    --qt;
    r += d;
  } else if (r >= d) {
    ++qt;
    r -= d;
  }
  rem = r;
  return qt;
}

// Cell-centre coordinate, calculate_p.py:48-50: start + (0.5 + i) * d.
__host__ __device__ inline double centre(double start, int i, double d) {
  return start + (0.5 + double(i)) * d;
}

}  // namespace bx
