// The nonlinear FVM stage in the q1-line layout [q2][p][q1] (q1 fastest), for
// 2-D or 1-D q-space with fields off -- the layout the solver picks when the
// q1 extent is a multiple of 64.
//
// qline_kernel: a wave owns one p-cell and the whole periodic q1 line (NR
// cells per lane, lane l holding q1 = l + 64 k), and marches along q2 over a
// segment of rows:
//  * q2: a 6-row register window per cell (five stencil rows and the row
//    loaded ahead); the WENO indicator and the upwind flux of each q2 face are
//    computed once and carried to the next row;
//  * q1: the row, its indicators and its face fluxes go through a
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
// Fused multiply-add of the march.  The march is compiled as several copies
// (one per register rotation), and the compiler's choice of which product of
// a*b + c*d to fuse depends on how often each product is reused, which
// differs between copies; so the march's functions evaluate with contraction
// off and fuse explicitly, and every row rounds the same way in every copy.
__host__ __device__ inline double ql_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__host__ __device__ inline float ql_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <int R, typename T>
__host__ __device__ inline T ql_kappa(T a, T b, T c) {
#pragma clang fp contract(off)
  if constexpr (R != RC_WENO5) {
    return T(0);
  } else if constexpr (sizeof(T) == 8) {
    const double bt = weno_beta_literal(a, b, c);
    return bt * bt;
  } else {
    // 13/12 (a - 2b + c)^2 + 1/4 (a - 4b + 3c)^2 + eps (1 + |a| + |b| + |c|)
    const float s1 = ql_fma(-2.0f, b, a) + c;
    const float s2 = ql_fma(3.0f, c, ql_fma(-4.0f, b, a));
    const float e = 1e-17f * (((1.0f + aabs(a)) + aabs(b)) + aabs(c));
    const float bt = ql_fma((13.0f / 12.0f) * s1, s1, ql_fma(0.25f * s2, s2, e));
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
#pragma clang fp contract(off)
  if constexpr (R == RC_WENO5) {
    T u1, u2, u3;
    if constexpr (RIGHT) {
      u1 = ql_fma(T(1.875), y2, ql_fma(T(-1.25), y1, T(0.375) * y0));
      u2 = ql_fma(T(3.75), y3, ql_fma(T(7.5), y2, T(-1.25) * y1));
      u3 = ql_fma(T(-0.625), y4, ql_fma(T(3.75), y3, T(1.875) * y2));
    } else {
      u1 = ql_fma(T(1.875), y2, ql_fma(T(3.75), y1, T(-0.625) * y0));
      u2 = ql_fma(T(-1.25), y3, ql_fma(T(7.5), y2, T(3.75) * y1));
      u3 = ql_fma(T(0.375), y4, ql_fma(T(-1.25), y3, T(1.875) * y2));
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
    const T num = ql_fma(w3, u3, ql_fma(w2, u2, w1 * u1));
    const T den = RIGHT ? ql_fma(T(5), w3, ql_fma(T(10), w2, w1))
                        : ql_fma(T(5), w1, ql_fma(T(10), w2, w3));
    return num * fast_rcp(den);
  } else {
    return face_value<R>(y0, y1, y2, y3, y4, RIGHT);
  }
}

// Upwind flux C * state through one face, rounded once (riemann.py:46-49 then
// the C_q product of df_dt_fvm.py:56-85): each face's flux is formed once and
// carried (q2) or shared through LDS (q1), so both cells of a face see the
// same rounded value, whatever code copy computed it.
template <typename T>
__device__ __forceinline__ T ql_flux(T C, T state) {
#pragma clang fp contract(off)
  return C * state;
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

// Every f-sized input stream (window rows of u, w, un) is read once per
// launch, so it goes through non-temporal loads and the L2 keeps the
// Maxwellian table rows the p-cells of a segment share; the output is stored
// normally, since the moment pass reads it right after (measured 1-2 % faster
// per step than a non-temporal store).
//
// One wave's march over a q2 segment for the upwind directions P1 (C_q1 > 0)
// and P2 (C_q2 > 0), fixed at compile time, in stage form MODE with PD
// Maxwellian factors per cell (0 = no source; -1 = decided at run time).  Along q2 the register window
// holds five rows y0..y4: rows r-2..r+2 (P2: the face r+1/2 is the right face
// of row r) or r-1..r+3 (not P2: it is the left face of row r+1); either way
// the face uses the window's centre y2 and the indicators of y1, y2, y3, and
// the flux through face r-1/2 is carried from the previous row.  Along q1 the
// row goes through the wave's LDS line: values, then indicators, then the flux
// through each cell's lower face.
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

  // Register roles rotate instead of sliding: window row j of the row being
  // updated lives in Y[(S + j) % 6], the indicators of its cells y1, y2 in
  // K[S % 3], K[(S + 1) % 3], the carried lower q2 face flux and the row's
  // pointwise inputs in slot S % 2, with S = it % 6.  Every load then lands in
  // the register that consumes it, so no copy of a loaded value forces a wait
  // on the loads (and stores) still in flight.
  T Y[6][NR], K[3][NR], S2[2][NR], WC[2][NR], UNC[2][NR], TC1[2][NR], TC2[2][NR], TC3[2][NR];
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
      for (int k = 0; k < NR; ++k) Y[j][k] = row[i1[k]];
    }
    const int rs0 = qm_mod(r0, n2);
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      WC[0][k] = w_is_u ? T(0) : wv[rs0 * rowP + i1[k]];
      UNC[0][k] = use_un ? unv[rs0 * rowP + i1[k]] : T(0);
      TC1[0][k] = src ? t1p[(long)rs0 * n1 + i1[k]] : T(0);
      TC2[0][k] = src2 ? t2p[(long)rs0 * n1 + i1[k]] : T(0);
      TC3[0][k] = src3 ? t3p[(long)rs0 * n1 + i1[k]] : T(0);
      S2[0][k] = T(0);
      K[0][k] = K[1][k] = T(0);
      if (do_q2) {
        const T k0 = ql_kappa<R, T>(z[k], Y[0][k], Y[1][k]);
        K[0][k] = ql_kappa<R, T>(Y[0][k], Y[1][k], Y[2][k]);
        K[1][k] = ql_kappa<R, T>(Y[1][k], Y[2][k], Y[3][k]);
        S2[0][k] = ql_flux(C2, ql_face<R, P2, T>(z[k], Y[0][k], Y[1][k], Y[2][k], Y[3][k], k0,
                                                 K[0][k], K[1][k]));
      }
    }
  }
  // LDS addresses each lane cell is written to: the cell and its periodic
  // halo copy (or the cell again); a lane past the line writes both to a
  // private slot past the halo, so every write is unconditional
  int a1[NR], h1[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    a1[k] = live[k] ? i1[k] : NR * 64 + QL_HALO + lane;
    h1[k] = !live[k] ? a1[k]
            : i1[k] < QL_HALO ? i1[k] + n1
            : i1[k] >= n1 - QL_HALO ? i1[k] - n1 : i1[k];
  }
  // periodic indices of the current row, the row after it and the window row
  // that enters the window at the next row
  int rcur = qm_mod(r0, n2);
  int rnx1 = rcur + 1 == n2 ? 0 : rcur + 1;
  int rnxw = qm_mod(r0 + B + 5, n2);

  auto row = [&](auto rot) {
#pragma clang fp contract(off)
    constexpr int S = decltype(rot)::value;
    constexpr int y0 = S % 6, y1 = (S + 1) % 6, y2 = (S + 2) % 6, y3 = (S + 3) % 6,
                  y4 = (S + 4) % 6, yl = (S + 5) % 6;
    constexpr int ka = S % 3, kb = (S + 1) % 3, kc = (S + 2) % 3;
    constexpr int cu = S % 2, nx = (S + 1) % 2;
    // the window row of the next row and the next row's pointwise inputs,
    // one row ahead, so the loads stay in flight across the stencil work
    {
      const T* uw = u + rnxw * rowP;
      const T* w1 = wv + rnx1 * rowP;
      const T* un1 = unv + rnx1 * rowP;
      const long tq = (long)rnx1 * n1;
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        Y[yl][k] = __builtin_nontemporal_load(uw + i1[k]);
        WC[nx][k] = w_is_u ? T(0) : __builtin_nontemporal_load(w1 + i1[k]);
        UNC[nx][k] = use_un ? __builtin_nontemporal_load(un1 + i1[k]) : T(0);
        TC1[nx][k] = src ? t1p[tq + i1[k]] : T(0);
        TC2[nx][k] = src2 ? t2p[tq + i1[k]] : T(0);
        TC3[nx][k] = src3 ? t3p[tq + i1[k]] : T(0);
      }
    }
    T d[NR], fc[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      fc[k] = P2 ? Y[y2][k] : Y[y1][k];
      d[k] = T(0);
    }
    if (do_q2) {
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        K[kc][k] = ql_kappa<R, T>(Y[y2][k], Y[y3][k], Y[y4][k]);
        const T f2p = ql_flux(C2, ql_face<R, P2, T>(Y[y0][k], Y[y1][k], Y[y2][k], Y[y3][k],
                                                    Y[y4][k], K[ka][k], K[kb][k], K[kc][k]));
        d[k] = (f2p - S2[cu][k]) * tinv2;
        S2[nx][k] = f2p;
      }
    }
    if (do_q1) {
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        fl[a1[k]] = fc[k];
        fl[h1[k]] = fc[k];
      }
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
        if constexpr (R == RC_WENO5) {
          kl[a1[k]] = kown[k];
          kl[h1[k]] = kown[k];
        }
      }
      __builtin_amdgcn_wave_barrier();
      T sf[NR];
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        if constexpr (P1) {
          // right face of cell i-1: window i-3..i+1, indicators of i-2, i-1, i
          T k2 = T(0), k1 = T(0);
          if constexpr (R == RC_WENO5) {
            k2 = kl[i1[k] - 2];
            k1 = kl[i1[k] - 1];
          }
          sf[k] = ql_face<R, true, T>(va[k], vb[k], vc[k], fc[k], vd[k], k2, k1, kown[k]);
        } else {
          // left face of cell i: window i-2..i+2, indicators of i-1, i, i+1
          T k1 = T(0), kp = T(0);
          if constexpr (R == RC_WENO5) {
            k1 = kl[i1[k] - 1];
            kp = kl[i1[k] + 1];
          }
          sf[k] = ql_face<R, false, T>(va[k], vb[k], fc[k], vc[k], vd[k], k1, kown[k], kp);
        }
        sf[k] = ql_flux(C1, sf[k]);
        sl[a1[k]] = sf[k];
        sl[h1[k]] = sf[k];
      }
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        const T fu = sl[i1[k] + 1];  // flux through the lower face of cell i+1
        d[k] = ql_fma(sf[k] - fu, tinv1, -d[k]);  // -(fu - sf) / dq1 - d
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
      const T f = fc[k];
      T dk = d[k];
      if (src) {
        T fmb = TC1[cu][k];
        if (src2) fmb = fmb * TC2[cu][k];
        if (src3) fmb = fmb * TC3[cu][k];
        dk = ql_fma(fmb - f, tinvtau, dk);
      }
      const T tv = ql_fma(dk, tcdt, w_is_u ? f : WC[cu][k]);
      const T x = use_un ? ql_fma(ta, UNC[cu][k], tb * tv) : (plain ? tv : tb * tv);
      if (live[k]) orow[i1[k]] = x;  // kept in the caches for the moment pass
    }
    rcur = rnx1;
    rnx1 = rnx1 + 1 == n2 ? 0 : rnx1 + 1;
    rnxw = rnxw + 1 == n2 ? 0 : rnxw + 1;
  };
  using std::integral_constant;
  // six rows per pass, so the rotation is back at 0 on the loop's back edge
  for (int it = 0; it < len;) {
    row(integral_constant<int, 0>());
    if (++it == len) break;
    row(integral_constant<int, 1>());
    if (++it == len) break;
    row(integral_constant<int, 2>());
    if (++it == len) break;
    row(integral_constant<int, 3>());
    if (++it == len) break;
    row(integral_constant<int, 4>());
    if (++it == len) break;
    row(integral_constant<int, 5>());
    ++it;
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
  __shared__ T fl[4][NR * 64 + 2 * QL_HALO + 64];
  __shared__ T kl[4][NR * 64 + 2 * QL_HALO + 64];
  __shared__ T sl[4][NR * 64 + 2 * QL_HALO + 64];
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
