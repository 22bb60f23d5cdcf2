// Fused RK stage, one workgroup per q-cell (all P of it), for any
// dimensionality, fields on or off.  Used with a workspace that holds the
// sign-grouped p order ("slabs"): each wave processes 64 p-cells whose
// velocities v1 and v2 have one sign each, so the upwind direction of the q1
// and q2 faces is uniform across the wave and its branch is scalar.
//
// Per phase-space cell and q axis the upwind states of both faces are built
// from four WENO indicators (each face needs three; they share two), and the
// upwind flux difference C*(s(i+1/2) - s(i-1/2)) is formed directly.  The
// local Lax-Friedrichs flux 0.5*(C(fL+fR) - |C|(fR-fL)) with alpha = |C| is
// the same function for this linear flux (C does not depend on f), so LLF
// evaluates through the upwind state as well.
//
// Reference: df_dt_fvm.py:56-97 (q fluxes + source), 251-323 (p fluxes),
// riemann.py:46-49, weno5.py:57-98, collision_operator.py:52-65.
#pragma once
#include "fvm_stage.h"

namespace bx {

// WENO5 face of the centre cell y2 from its three indicators, FP64: the
// normalised weights d_k/b_k^2 (d = 1, 10, 5 for the right face and 5, 10, 1
// for the left face, weno5.py:73-80 scaled by 16) applied through one division,
//   sum_k d_k u_k prod_{j!=k} b_j^2 / sum_k d_k prod_{j!=k} b_j^2.
// (b >= 1e-17, so the products stay far inside the FP64 range.)
__host__ __device__ inline double weno_face_betas(double y0, double y1, double y2, double y3, double y4,
                                         double b1, double b2, double b3, bool right) {
  const double s1 = b1 * b1, s2 = b2 * b2, s3 = b3 * b3;
  const double p23 = s2 * s3, p13 = s1 * s3, p12 = s1 * s2;
  double t1, t2, t3, u1, u2, u3;
  if (right) {
    t1 = p23;
    t2 = 10.0 * p13;
    t3 = 5.0 * p12;
    u1 = 0.375 * y0 - 1.25 * y1 + 1.875 * y2;
    u2 = -0.125 * y1 + 0.75 * y2 + 0.375 * y3;
    u3 = 0.375 * y2 + 0.75 * y3 - 0.125 * y4;
  } else {
    t1 = 5.0 * p23;
    t2 = 10.0 * p13;
    t3 = p12;
    u1 = -0.125 * y0 + 0.75 * y1 + 0.375 * y2;
    u2 = 0.375 * y1 + 0.75 * y2 - 0.125 * y3;
    u3 = 1.875 * y2 - 1.25 * y3 + 0.375 * y4;
  }
  return (t1 * u1 + t2 * u2 + t3 * u3) * fast_rcp(t1 + t2 + t3);
}

// FP32: the squares of tiny indicators would underflow in products, so the
// weights are formed as reciprocals (weno5.py:73-96).
__host__ __device__ inline float weno_face_betas(float y0, float y1, float y2, float y3, float y4, float b1,
                                        float b2, float b3, bool right) {
  return weno_combine(y0, y1, y2, y3, y4, b1, b2, b3, right);
}

// Upwind flux difference C*(s(i+1/2) - s(i-1/2)) of cell i from a[k] =
// f[i-3+k] with the upwind direction ``pos`` (C > 0) uniform over the wave;
// the upwind state of the lower face s(i-1/2) (f_left / f_bot of
// df_dt_fvm.py:79-80) is stored in ``lower``.
template <int R, typename T>
__host__ __device__ inline T upwind_flux_diff(const T a[7], T C, bool pos, T& lower) {
  T sL, sR;
  if constexpr (R == RC_WENO5) {
    if (pos) {
      // right faces of cells i-1 and i: windows a0..a4 and a1..a5
      const T b0 = weno_beta<T>(a[0], a[1], a[2]), b1 = weno_beta<T>(a[1], a[2], a[3]);
      const T b2 = weno_beta<T>(a[2], a[3], a[4]), b3 = weno_beta<T>(a[3], a[4], a[5]);
      sL = weno_face_betas(a[0], a[1], a[2], a[3], a[4], b0, b1, b2, true);
      sR = weno_face_betas(a[1], a[2], a[3], a[4], a[5], b1, b2, b3, true);
    } else {
      // left faces of cells i and i+1: windows a1..a5 and a2..a6
      const T b1 = weno_beta<T>(a[1], a[2], a[3]), b2 = weno_beta<T>(a[2], a[3], a[4]);
      const T b3 = weno_beta<T>(a[3], a[4], a[5]), b4 = weno_beta<T>(a[4], a[5], a[6]);
      sL = weno_face_betas(a[1], a[2], a[3], a[4], a[5], b1, b2, b3, false);
      sR = weno_face_betas(a[2], a[3], a[4], a[5], a[6], b2, b3, b4, false);
    }
  } else {
    if (pos) {
      sL = face_value<R>(a[0], a[1], a[2], a[3], a[4], true);
      sR = face_value<R>(a[1], a[2], a[3], a[4], a[5], true);
    } else {
      sL = face_value<R>(a[1], a[2], a[3], a[4], a[5], false);
      sR = face_value<R>(a[2], a[3], a[4], a[5], a[6], false);
    }
  }
  lower = sL;
  return C * sR - C * sL;
}

template <int R, typename T>
__host__ __device__ inline T upwind_flux_diff(const T a[7], T C, bool pos) {
  T lower;
  return upwind_flux_diff<R, T>(a, C, pos, lower);
}

// Flux difference along a p axis of the q-cell's p-block ``blk`` (stride
// ``stride``, n cells, periodic), from the upwind states only.  C_p1 =
// (q/m)(E1 + v2 B3 - v3 B2) does not depend on v1 (advection_terms.py:147-188),
// and likewise for p2 and p3, so C is the same on both sides of every face
// along its own axis and LLF = upwind for this linear flux.
template <int R, typename T>
__device__ inline T p_upwind_flux_diff(const T* blk, int p, int j, int n, int stride, T C) {
  T a[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    int jj = j + k - 3;
    if (n >= 4) {
      jj = jj < 0 ? jj + n : (jj >= n ? jj - n : jj);
    } else {
      jj = ((jj % n) + n) % n;
    }
    a[k] = blk[p + (jj - j) * stride];
  }
  return upwind_flux_diff<R, T>(a, C, C > T(0));
}

struct QCParams {
  StageParams sp;
  const int* perm;    // [nslab*64] p of each slab lane, -1 = empty lane
  const int* jpack;   // [nslab*64] j1 | j2 << 10 | j3 << 20 of the same p
  int nslab;
};

// Workgroup size (threads NT, a template parameter): with p-advection the
// whole p-block of the q-cell sits in LDS.  A block that fits twice into the
// 160 KB of a CU (FP32 at 24^3 p-cells: 55 KB) runs as two workgroups of 8
// waves, so one fills its block while the other computes; a larger one (FP64
// at 24^3: 110 KB) as one workgroup of 16 waves (the launch picks NT).
// Without p-advection the workgroup is 4 waves.

// q-cell of workgroup b of nb = Q: the eight XCDs take contiguous bands of
// q2 rows, and each walks its band q2-fastest, so a q-cell's stencil
// neighbours (3 rows and 3 columns away) were loaded within the last few
// dozen workgroups of that XCD and are still in its L2.
__device__ inline int qc_order(const StageParams& sp, int b, int nb) {
  if (nb % 8 != 0) return b;
  const int l = b / 8, x = b % 8;
  if (sp.n_q2 % 8 != 0 || sp.n_q2 < 8) return x * (nb / 8) + l;
  const int rows = sp.n_q2 / 8;  // rows of the band
  return (l / rows) + sp.n_q1 * (x * rows + l % rows);
}

// End of a slab kernel: the workgroup sums (butterfly within waves, waves in
// order) of the five moments of the output and, if requested, the three
// currents of the input, written as the q-cell's moment block and J.
template <int NW>
__device__ inline void qc_finish(const StageParams& sp, int q, double (&acc)[5], double (&jac)[3],
                                 double (*red)[NW], int wave, int lane) {
  const bool mom = sp.mom_out != nullptr, cur = sp.Jout != nullptr;
  if (!mom && !cur) return;
  double v[8];
#pragma unroll
  for (int k = 0; k < 5; ++k) v[k] = acc[k];
#pragma unroll
  for (int k = 0; k < 3; ++k) v[5 + k] = jac[k];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    double x = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    v[k] = x;
  }
  if (lane == 0)
    for (int k = 0; k < 8; ++k) red[k][wave] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    double m[8];
    for (int k = 0; k < 8; ++k) {
      double x = red[k][0];
      for (int w = 1; w < NW; ++w) x += red[k][w];
      m[k] = x * sp.measure;
    }
    if (mom) write_moment_block(sp.mom_out, sp.Q, q, m, sp.mass, sp.kB, sp.p_dim);
    if (cur)
      for (int k = 0; k < 3; ++k) sp.Jout[k * sp.Q + q] = sp.charge * m[5 + k];
  }
}

template <typename T, int RQ, int RP, bool PADV, int NT>
__global__ __launch_bounds__(NT) void qcell_kernel(QCParams qp) {
  constexpr int NW = NT / 64;
  extern __shared__ unsigned char smem[];
  __shared__ double red[8][NW];
  const StageParams& sp = qp.sp;
  T* tab = reinterpret_cast<T*>(smem);
  T* blk = tab + (sp.n_p1 + sp.n_p2 + sp.n_p3);
  const int q = qc_order(sp, blockIdx.x, gridDim.x);
  QCell<T> c;
  qcell_neighbours(sp, q, c);
  c.pref = 0.0;
  c.e1 = c.e2 = c.e3 = tab;
  if (sp.source) {
    const BgkCell b = qcell_bgk(sp, q);
    c.pref = maxwell_prefactor(b, sp.mass, sp.kB, sp.p_dim);
    T* e1 = tab;
    T* e2 = tab + sp.n_p1;
    T* e3 = e2 + sp.n_p2;
    for (int i = threadIdx.x; i < sp.n_p1; i += blockDim.x)
      e1[i] = T(maxwell_factor(centre(sp.p1s, i, sp.dp1), b.u1, b.T, sp.mass, sp.kB));
    for (int i = threadIdx.x; i < sp.n_p2; i += blockDim.x)
      e2[i] = T(maxwell_factor(centre(sp.p2s, i, sp.dp2), b.u2, b.T, sp.mass, sp.kB));
    for (int i = threadIdx.x; i < sp.n_p3; i += blockDim.x)
      e3[i] = T(maxwell_factor(centre(sp.p3s, i, sp.dp3), b.u3, b.T, sp.mass, sp.kB));
    c.e1 = e1;
    c.e2 = e2;
    c.e3 = e3;
  }
  const T* u = (const T*)sp.u;
  const long P = sp.P;
  if constexpr (PADV) {
    qcell_fields(sp, c);
    for (int p = threadIdx.x; p < sp.P; p += blockDim.x) blk[p] = u[(long)q * P + p];
  }
  __syncthreads();

  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const T* wv = (const T*)sp.w;
  const T* unv = (const T*)sp.un;
  T* outv = (T*)sp.out;
  const bool use_un = sp.a != 0.0;
  const bool w_is_u = sp.w == sp.u;
  const T ta = T(sp.a), tb = T(sp.b), tcdt = T(sp.cdt);
  const T tinv1 = T(sp.inv_dq1), tinv2 = T(sp.inv_dq2), tinvtau = T(sp.inv_tau);
  constexpr int r = recon_reach(RQ) + 1;
  const long base = (long)q * P;
  long nb1[7], nb2[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    nb1[k] = (long)c.nb1[k] * P;
    nb2[k] = (long)c.nb2[k] * P;
  }
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  double jac[3] = {0.0, 0.0, 0.0};
  for (int s = wave; s < qp.nslab; s += NW) {
    const int p_lane = qp.perm[s * 64 + lane];
    const bool valid = p_lane >= 0;
    const int p = valid ? p_lane : 0;
    const int jp = valid ? qp.jpack[s * 64 + lane] : 0;
    const int j1 = jp & 1023, j2 = (jp >> 10) & 1023, j3 = jp >> 20;
    const double v1 = centre(sp.p1s, j1, sp.dp1);
    const double v2 = centre(sp.p2s, j2, sp.dp2);
    const double v3 = centre(sp.p3s, j3, sp.dp3);
    const bool pos1 = __builtin_amdgcn_readfirstlane(int(v1 > 0.0)) != 0;
    const bool pos2 = __builtin_amdgcn_readfirstlane(int(v2 > 0.0)) != 0;
    const T f = u[base + p];
    T t1 = T(0), t2 = T(0), s1 = f, s2 = f;
    if (sp.n_q1 > 1) {
      T a[7];
#pragma unroll
      for (int k = 3 - r; k <= 3 + r; ++k) a[k] = k == 3 ? f : u[nb1[k] + p];
      t1 = upwind_flux_diff<RQ, T>(a, T(v1), pos1, s1);
    }
    if (sp.n_q2 > 1) {
      T a[7];
#pragma unroll
      for (int k = 3 - r; k <= 3 + r; ++k) a[k] = k == 3 ? f : u[nb2[k] + p];
      t2 = upwind_flux_diff<RQ, T>(a, T(v2), pos2, s2);
    }
    if (valid && sp.Jout) {
      jac[0] += double(s1) * v1;
      jac[1] += double(s2) * v2;
      jac[2] += double(f) * v3;
    }
    T d = -t1 * tinv1 - t2 * tinv2;
    if (sp.source) {
      T fmb = T(c.pref) * c.e1[j1];
      if (sp.p_dim >= 2) fmb = fmb * c.e2[j2];
      if (sp.p_dim >= 3) fmb = fmb * c.e3[j3];
      d = d + (fmb - f) * tinvtau;
    }
    if constexpr (PADV) {
      const double em = sp.charge / sp.mass;
      T dp = T(0);
      if (sp.n_p1 > 1) {
        const double C = em * (c.E[0] + v2 * c.B[0][2] - v3 * c.B[0][1]);
        dp = p_upwind_flux_diff<RP, T>(blk, p, j1, sp.n_p1, 1, T(C)) * T(sp.inv_dp1);
      }
      if (sp.n_p2 > 1) {
        const double C = em * (c.E[1] + v3 * c.B[1][0] - v1 * c.B[1][2]);
        dp = dp + p_upwind_flux_diff<RP, T>(blk, p, j2, sp.n_p2, sp.n_p1, T(C)) * T(sp.inv_dp2);
      }
      if (sp.n_p3 > 1) {
        const double C = em * (c.E[2] + v1 * c.B[2][1] - v2 * c.B[2][0]);
        dp = dp + p_upwind_flux_diff<RP, T>(blk, p, j3, sp.n_p3, sp.n_p1 * sp.n_p2, T(C)) *
                  T(sp.inv_dp3);
      }
      d = d - dp;
    }
    const T wval = w_is_u ? f : wv[base + p];
    const T tv = wval + d * tcdt;
    const T x = use_un ? ta * unv[base + p] + tb * tv : (sp.b == 1.0 ? tv : tb * tv);
    if (valid) {
      outv[base + p] = x;
      const double xd = double(x);
      acc[0] += xd;
      acc[1] += xd * v1;
      acc[2] += xd * v2;
      acc[3] += xd * v3;
      acc[4] += (0.5 * xd) * ((v1 * v1 + v2 * v2) + v3 * v3);
    }
  }
  qc_finish<NW>(sp, q, acc, jac, red, wave, lane);
}

}  // namespace bx
