// Ensembles of independent instances that are one-dimensional in q (periodic
// q1, N_q2 = 1, fields off), each with its own grid, velocity grid, collision
// time and time step, advanced together.  Every instance keeps f in the
// q1-line layout [P][n1] inside pooled arrays.
//
// One SSP-RK3 / RK2 stage of the whole ensemble is three launches:
//  * ens_stage_kernel: a persistent grid of workgroups, each walking its own
//    range of wave tasks (instance, p-range); a wave holds one or more
//    periodic q1 lines (n1 <= 64: 64/n1 lines per wave; larger n1: n1/64
//    cells per lane), exchanges the line through LDS and evaluates the
//    reconstruction, upwind flux, BGK source and RK combination of every
//    cell (df_dt_fvm.py:56-97, collision_operator.py:70-128,
//    fvm_operator.py:22-33);
//  * ens_moments_kernel: one workgroup per (instance, p-chunk), FP64
//    partial moments in a fixed order;
//  * ens_finalize_kernel: one workgroup per instance: the moment block
//    (moments.py:5-22, collision_operator.py:102-116) and the separable
//    Maxwellian tables the next stage reads.
// The work lists are built on the host (tile sizes from the tuning
// parameters, workgroup ranges balanced by cell count), so no kernel waits
// on another workgroup.
#pragma once
#include "qline.h"

namespace bx {

struct EnsInst {
  int n1, P, n_p1, n_p2, n_p3, p_dim, source, lanes;  // lanes: q1 cells per lane-group
  long off_f;     // element offset of the [P][n1] block in the pooled f-sized arrays
  long off_tab;   // element offset of the [n_p1+n_p2+n_p3][n1] Maxwellian tables
  long off_mom;   // double offset of the [12][n1] moment block
  long off_part;  // double offset of the [nchunk][5][n1] partial moments
  int nchunk, pad;
  double inv_dq, cdt, inv_tau, p1s, dp1, p2s, dp2, p3s, dp3, measure, mass, kB;
};

struct EnsTile {
  int inst, p0, np, pad;  // wave task: p-cells p0 .. p0+np-1 of instance inst
};

struct EnsChunk {
  int inst, chunk, p0, p1;  // moment workgroup: p-cells p0 .. p1-1
};

struct EnsStageArgs {
  const EnsInst* inst;
  const EnsTile* tiles;
  const int* wg_begin;  // tiles of workgroup g: [wg_begin[g], wg_begin[g+1])
  const void* u;
  const void* w;
  const void* un;
  void* out;
  const void* tab;
  double a, b, cscale;  // stage: out = a*un + b*(w + cscale*dt_i*L(u))
  int use_un, b_one;
};

// LDS per wave: q1 lines with QL_HALO periodic copies on each side; a wave
// packs 64 / n1 lines when n1 <= 64 (at most 12 lines of n1 + 6 cells for
// n1 >= 5) or holds one line of up to 256 cells.
constexpr int ENS_LDS = 264;

// Stencil, source and RK combination of the q1 lines of one wave.  Lane l
// works on cells i1 = c0 + 64 k (k < NR) of the line starting at ``line``
// (LDS, halo-padded; p-cell p), all lanes of a line sharing C1 and the table
// rows.
template <typename T, int R, int NR>
__device__ inline void ens_lines(const EnsStageArgs& ea, const EnsInst& I, int p, int lane_in_line,
                                 bool live_line, T* fl, T* kl, T* sl) {
  const int n1 = I.n1;
  const int j1 = p % I.n_p1, j23 = p / I.n_p1, j2 = j23 % I.n_p2, j3 = j23 / I.n_p2;
  const T C1 = T(centre(I.p1s, j1, I.dp1));
  const bool pos = C1 > T(0);
  const long base = I.off_f + (long)p * n1;
  const T* u = (const T*)ea.u + base;
  int i1[NR];
  bool live[NR];
  T f[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    i1[k] = lane_in_line + 64 * k;
    live[k] = live_line && i1[k] < n1;
    if (!live[k]) i1[k] = 0;
    f[k] = live[k] ? u[i1[k]] : T(0);
    ql_put(fl, i1[k], live[k], n1, f[k]);
  }
  __builtin_amdgcn_wave_barrier();
  T m3[NR], m2[NR], m1[NR], p1[NR], p2[NR], kap[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const T* line = fl + i1[k];
    m3[k] = line[-3];
    m2[k] = line[-2];
    m1[k] = line[-1];
    p1[k] = line[1];
    p2[k] = line[2];
    kap[k] = ql_kappa<R, T>(m1[k], f[k], p1[k]);
    ql_put(kl, i1[k], live[k], n1, kap[k]);
  }
  __builtin_amdgcn_wave_barrier();
  T s[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    T km2 = T(0), km1 = T(0), kp1 = T(0);
    if constexpr (R == RC_WENO5) {
      km2 = kl[i1[k] - 2];
      km1 = kl[i1[k] - 1];
      kp1 = kl[i1[k] + 1];
    }
    // lower face of cell i: right face of i-1 (C > 0) or left face of i
    s[k] = pos ? ql_face<R, true, T>(m3[k], m2[k], m1[k], f[k], p1[k], km2, km1, kap[k])
               : ql_face<R, false, T>(m2[k], m1[k], f[k], p1[k], p2[k], km1, kap[k], kp1);
    ql_put(sl, i1[k], live[k], n1, s[k]);
  }
  __builtin_amdgcn_wave_barrier();
  const T* tab = (const T*)ea.tab + I.off_tab;
  const T tinv = T(I.inv_dq), ttau = T(I.inv_tau), tcdt = T(I.cdt * ea.cscale);
  const T ta = T(ea.a), tb = T(ea.b);
  T* out = (T*)ea.out + base;
  const T* wv = (const T*)ea.w + base;
  const T* unv = (const T*)ea.un + base;
  const bool w_is_u = ea.w == ea.u;
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    if (!live[k]) continue;
    const int i = i1[k];
    const T su = sl[i + 1];
    T d = -(C1 * su - C1 * s[k]) * tinv;
    if (I.source) {
      T fmb = tab[(long)j1 * n1 + i];
      if (I.p_dim >= 2) fmb = fmb * tab[(long)(I.n_p1 + j2) * n1 + i];
      if (I.p_dim >= 3) fmb = fmb * tab[(long)(I.n_p1 + I.n_p2 + j3) * n1 + i];
      d = d + (fmb - f[k]) * ttau;
    }
    const T tv = (w_is_u ? f[k] : wv[i]) + d * tcdt;
    out[i] = ea.use_un ? ta * unv[i] + tb * tv : (ea.b_one ? tv : tb * tv);
  }
  __builtin_amdgcn_wave_barrier();
}

#ifndef BX_NO_KERNELS
// Persistent stage kernel: workgroup g walks the wave tasks
// [wg_begin[g], wg_begin[g+1]) with its WPG waves taking every WPG-th task.
template <typename T, int R>
__global__ __launch_bounds__(256) void ens_stage_kernel(EnsStageArgs ea) {
  __shared__ T fl[4][ENS_LDS];
  __shared__ T kl[4][ENS_LDS];
  __shared__ T sl[4][ENS_LDS];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int nwave = blockDim.x >> 6;
  const int tb = ea.wg_begin[blockIdx.x], te = ea.wg_begin[blockIdx.x + 1];
  for (int t = tb + wave; t < te; t += nwave) {
    const EnsTile tile = ea.tiles[t];
    const EnsInst I = ea.inst[tile.inst];
    if (I.n1 <= 64) {
      // 64 / n1 lines per wave, one cell per lane; line l starts at
      // l * (n1 + 2 halo) + halo in LDS
      const int lines = 64 / I.n1;
      const int line = lane / I.n1, c = lane % I.n1;
      const int off = (line < lines ? line : 0) * (I.n1 + 2 * QL_HALO) + QL_HALO;
      for (int q = 0; q < tile.np; q += lines) {
        const int p = tile.p0 + q + line;
        const bool live = line < lines && q + line < tile.np;
        ens_lines<T, R, 1>(ea, I, live ? p : tile.p0, c, live, fl[wave] + off, kl[wave] + off,
                           sl[wave] + off);
      }
    } else if (I.n1 <= 128) {
      for (int q = 0; q < tile.np; ++q)
        ens_lines<T, R, 2>(ea, I, tile.p0 + q, lane, true, fl[wave] + QL_HALO,
                           kl[wave] + QL_HALO, sl[wave] + QL_HALO);
    } else {
      for (int q = 0; q < tile.np; ++q)
        ens_lines<T, R, 4>(ea, I, tile.p0 + q, lane, true, fl[wave] + QL_HALO,
                           kl[wave] + QL_HALO, sl[wave] + QL_HALO);
    }
  }
}

// Partial moments of one (instance, p-chunk): thread t sums q1 cell t % n1
// over every (256/n1)-th p-cell of the chunk; the p-groups are added in a
// fixed order.
template <typename T>
__global__ __launch_bounds__(256) void ens_moments_kernel(const EnsInst* inst,
                                                          const EnsChunk* chunks, const T* f,
                                                          double* part) {
  __shared__ double red[5][256];
  const EnsChunk ch = chunks[blockIdx.x];
  const EnsInst I = inst[ch.inst];
  const int t = threadIdx.x, n1 = I.n1;
  const int groups = 256 / n1;
  const int q = t % n1, g = t / n1;
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0;
  if (g < groups) {
    const T* col = f + I.off_f + q;
    for (int p = ch.p0 + g; p < ch.p1; p += groups) {
      const int j1 = p % I.n_p1, j23 = p / I.n_p1, j2 = j23 % I.n_p2, j3 = j23 / I.n_p2;
      const double v1 = centre(I.p1s, j1, I.dp1), v2 = centre(I.p2s, j2, I.dp2);
      const double v3 = centre(I.p3s, j3, I.dp3);
      const double x = double(col[(long)p * n1]);
      a0 += x;
      a1 += x * v1;
      a2 += x * v2;
      a3 += x * v3;
      a4 += x * (0.5 * ((v1 * v1 + v2 * v2) + v3 * v3));
    }
  }
  red[0][t] = a0;
  red[1][t] = a1;
  red[2][t] = a2;
  red[3][t] = a3;
  red[4][t] = a4;
  __syncthreads();
  if (t < n1) {
    for (int k = 0; k < 5; ++k) {
      double s = 0.0;
      for (int gg = 0; gg < groups; ++gg) s += red[k][gg * n1 + t];
      part[I.off_part + ((long)ch.chunk * 5 + k) * n1 + t] = s;
    }
  }
}

// Moment block and Maxwellian tables of every instance (one workgroup per
// instance, one thread per q1 cell); a non-finite density sets *flag.
template <typename T>
__global__ __launch_bounds__(256) void ens_finalize_kernel(const EnsInst* inst,
                                                           const double* part, double* mom,
                                                           T* tab, int* flag) {
  const EnsInst I = inst[blockIdx.x];
  const int q = threadIdx.x;
  if (q >= I.n1) return;
  const int n1 = I.n1;
  double m[5];
  for (int k = 0; k < 5; ++k) {
    double s = 0.0;
    for (int c = 0; c < I.nchunk; ++c) s += part[I.off_part + ((long)c * 5 + k) * n1 + q];
    m[k] = s * I.measure;
  }
  if (!isfinite(m[0])) *flag = 1;
  double* mb = mom + I.off_mom;
  write_moment_block(mb, n1, q, m, I.mass, I.kB, I.p_dim);
  if (!I.source) return;
  const double* b = mb + MOM_BGK * n1 + q;
  const double Tq = b[4 * n1];
  T* tb = tab + I.off_tab;
  for (int j = 0; j < I.n_p1; ++j)
    tb[(long)j * n1 + q] = T(b[0] * maxwell_factor(centre(I.p1s, j, I.dp1), b[n1], Tq, I.mass, I.kB));
  for (int j = 0; j < I.n_p2; ++j)
    tb[(long)(I.n_p1 + j) * n1 + q] =
        T(maxwell_factor(centre(I.p2s, j, I.dp2), b[2 * n1], Tq, I.mass, I.kB));
  for (int j = 0; j < I.n_p3; ++j)
    tb[(long)(I.n_p1 + I.n_p2 + j) * n1 + q] =
        T(maxwell_factor(centre(I.p3s, j, I.dp3), b[3 * n1], Tq, I.mass, I.kB));
}
#endif

}  // namespace bx
