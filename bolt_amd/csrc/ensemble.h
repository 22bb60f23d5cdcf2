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
#include <type_traits>
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

// Everything the stage reads from memory for one line of a wave (the lane's
// cells of it): u, the Maxwellian table factors, w and un, and the line's C1.
// A wave fetches the next line into one of two such slots while it works on
// the other, and the slots swap roles from line to line, so no loaded value
// is copied (a copy would wait for the loads still in flight).
template <typename T, int NR>
struct EnsLineIn {
  T f[NR], t1[NR], t2[NR], t3[NR], wc[NR], unc[NR];
  T C1;
};

// Fetch line p (velocity indices j1, j2, j3) for lane cells
// lane_in_line + 64 k; nothing is loaded for dead lanes or a dead line.
template <typename T, int NR>
__device__ inline void ens_fetch(const EnsStageArgs& ea, const EnsInst& I, int p, int j1, int j2,
                                 int j3, int lane_in_line, bool live_line, EnsLineIn<T, NR>& in) {
  const int n1 = I.n1;
  const long base = I.off_f + (long)p * n1;
  const T* u = (const T*)ea.u + base;
  const T* tab = (const T*)ea.tab + I.off_tab;
  const T* wv = (const T*)ea.w + base;
  const T* unv = (const T*)ea.un + base;
  const bool w_is_u = ea.w == ea.u;
  const bool src = I.source != 0, src2 = src && I.p_dim >= 2, src3 = src && I.p_dim >= 3;
  in.C1 = T(centre(I.p1s, j1, I.dp1));
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const int i = lane_in_line + 64 * k;
    const bool live = live_line && i < n1;
    in.f[k] = live ? __builtin_nontemporal_load(u + i) : T(0);
    in.t1[k] = live && src ? tab[(long)j1 * n1 + i] : T(0);
    in.t2[k] = live && src2 ? tab[(long)(I.n_p1 + j2) * n1 + i] : T(0);
    in.t3[k] = live && src3 ? tab[(long)(I.n_p1 + I.n_p2 + j3) * n1 + i] : T(0);
    in.wc[k] = live && !w_is_u ? __builtin_nontemporal_load(wv + i) : T(0);
    in.unc[k] = live && ea.use_un ? __builtin_nontemporal_load(unv + i) : T(0);
  }
}

// Stencil, source and RK combination of the q1 lines of one wave.  Lane l
// works on cells i1 = c0 + 64 k (k < NR) of the line starting at ``line``
// (LDS, halo-padded; p-cell p), all lanes of a line sharing C1 and the table
// rows.
template <typename T, int R, int NR>
__device__ inline void ens_lines(const EnsStageArgs& ea, const EnsInst& I, int p, int lane_in_line,
                                 bool live_line, const EnsLineIn<T, NR>& in, T* fl, T* kl, T* sl) {
  const int n1 = I.n1;
  const T C1 = in.C1;
  const bool pos = C1 > T(0);
  const long base = I.off_f + (long)p * n1;
  const bool w_is_u = ea.w == ea.u;
  int i1[NR];
  bool live[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    i1[k] = lane_in_line + 64 * k;
    live[k] = live_line && i1[k] < n1;
    if (!live[k]) i1[k] = 0;
  }
#pragma unroll
  for (int k = 0; k < NR; ++k) ql_put(fl, i1[k], live[k], n1, in.f[k]);
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
    kap[k] = ql_kappa<R, T>(m1[k], in.f[k], p1[k]);
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
    s[k] = pos ? ql_face<R, true, T>(m3[k], m2[k], m1[k], in.f[k], p1[k], km2, km1, kap[k])
               : ql_face<R, false, T>(m2[k], m1[k], in.f[k], p1[k], p2[k], km1, kap[k], kp1);
    ql_put(sl, i1[k], live[k], n1, s[k]);
  }
  __builtin_amdgcn_wave_barrier();
  const T tinv = T(I.inv_dq), ttau = T(I.inv_tau), tcdt = T(I.cdt * ea.cscale);
  const T ta = T(ea.a), tb = T(ea.b);
  T* out = (T*)ea.out + base;
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    if (!live[k]) continue;
    const int i = i1[k];
    const T f = in.f[k];
    const T su = sl[i + 1];
    T d = -(C1 * su - C1 * s[k]) * tinv;
    if (I.source) {
      T fmb = in.t1[k];
      if (I.p_dim >= 2) fmb = fmb * in.t2[k];
      if (I.p_dim >= 3) fmb = fmb * in.t3[k];
      d = d + (fmb - f) * ttau;
    }
    const T tv = (w_is_u ? f : in.wc[k]) + d * tcdt;
    __builtin_nontemporal_store(ea.use_un ? ta * in.unc[k] + tb * tv : (ea.b_one ? tv : tb * tv),
                                out + i);
  }
  __builtin_amdgcn_wave_barrier();
}

// (j1, j2, j3) of the p-cell ``step`` further on.
__device__ inline void ens_advance(const EnsInst& I, int step, int& j1, int& j2, int& j3) {
  j1 += step;
  while (j1 >= I.n_p1) {
    j1 -= I.n_p1;
    if (++j2 == I.n_p2) {
      j2 = 0;
      ++j3;
    }
  }
}

// A wave task of lines longer than 64 cells: one line at a time, the next
// line's inputs fetched while the current one is worked on.
template <typename T, int R, int NR>
__device__ inline void ens_task(const EnsStageArgs& ea, const EnsInst& I, const EnsTile& tile,
                                int lane, T* fl, T* kl, T* sl) {
  // indices of the next line to fetch
  int j1 = tile.p0 % I.n_p1, j23 = tile.p0 / I.n_p1;
  int j2 = j23 % I.n_p2, j3 = j23 / I.n_p2;
  EnsLineIn<T, NR> A, B;
  ens_fetch<T, NR>(ea, I, tile.p0, j1, j2, j3, lane, true, A);
  ens_advance(I, 1, j1, j2, j3);
  for (int q = 0; q < tile.np;) {
    bool more = q + 1 < tile.np;
    ens_fetch<T, NR>(ea, I, more ? tile.p0 + q + 1 : tile.p0, j1, j2, j3, lane, more, B);
    ens_advance(I, 1, j1, j2, j3);
    ens_lines<T, R, NR>(ea, I, tile.p0 + q, lane, true, A, fl, kl, sl);
    if (++q == tile.np) break;
    more = q + 1 < tile.np;
    ens_fetch<T, NR>(ea, I, more ? tile.p0 + q + 1 : tile.p0, j1, j2, j3, lane, more, A);
    ens_advance(I, 1, j1, j2, j3);
    ens_lines<T, R, NR>(ea, I, tile.p0 + q, lane, true, B, fl, kl, sl);
    ++q;
  }
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
      // p-cell of this lane's line and its (j1, j2, j3), carried without
      // divisions as the task advances by ``lines`` p-cells; the two input
      // slots alternate as in ens_task
      const bool inline_ = line < lines;
      const int p_first = tile.p0 + (inline_ ? line : 0);
      int j1 = p_first % I.n_p1, j23 = p_first / I.n_p1;
      int j2 = j23 % I.n_p2, j3 = j23 / I.n_p2;
      T* f_ = fl[wave] + off;
      T* k_ = kl[wave] + off;
      T* s_ = sl[wave] + off;
      EnsLineIn<T, 1> A, B;
      ens_fetch<T, 1>(ea, I, p_first, j1, j2, j3, c, inline_ && line < tile.np, A);
      ens_advance(I, lines, j1, j2, j3);
      for (int q = 0; q < tile.np;) {
        bool live = inline_ && q + line < tile.np;
        bool more = inline_ && q + lines + line < tile.np;
        ens_fetch<T, 1>(ea, I, more ? tile.p0 + q + lines + line : tile.p0, j1, j2, j3, c, more, B);
        ens_advance(I, lines, j1, j2, j3);
        ens_lines<T, R, 1>(ea, I, live ? tile.p0 + q + line : tile.p0, c, live, A, f_, k_, s_);
        q += lines;
        if (q >= tile.np) break;
        live = inline_ && q + line < tile.np;
        more = inline_ && q + lines + line < tile.np;
        ens_fetch<T, 1>(ea, I, more ? tile.p0 + q + lines + line : tile.p0, j1, j2, j3, c, more, A);
        ens_advance(I, lines, j1, j2, j3);
        ens_lines<T, R, 1>(ea, I, live ? tile.p0 + q + line : tile.p0, c, live, B, f_, k_, s_);
        q += lines;
      }
    } else if (I.n1 <= 128) {
      ens_task<T, R, 2>(ea, I, tile, lane, fl[wave] + QL_HALO, kl[wave] + QL_HALO,
                        sl[wave] + QL_HALO);
    } else {
      ens_task<T, R, 4>(ea, I, tile, lane, fl[wave] + QL_HALO, kl[wave] + QL_HALO,
                        sl[wave] + QL_HALO);
    }
  }
}

// Partial moments of one (instance, p-chunk): the chunk is a contiguous
// span of (p1 - p0) * n1 elements, streamed in 16-byte vectors (non-temporal,
// four in flight per thread) when n1 is a multiple of the vector width, each thread
// keeping the same VEC q1 cells; otherwise one element per thread.  The
// partial sums of threads holding the same cells are added in a fixed order.
template <typename T>
__global__ __launch_bounds__(256) void ens_moments_kernel(const EnsInst* inst,
                                                          const EnsChunk* chunks, const T* f,
                                                          double* part) {
  constexpr int VEC = 16 / sizeof(T);
  typedef T V __attribute__((ext_vector_type(VEC)));
  __shared__ double red[5][256];
  const EnsChunk ch = chunks[blockIdx.x];
  const EnsInst I = inst[ch.inst];
  const int t = threadIdx.x, n1 = I.n1;
  const bool vec = n1 % VEC == 0 && I.off_f % VEC == 0;  // 16-byte aligned rows
  const int w = vec ? VEC : 1;          // cells per thread
  const int lpl = n1 / w;               // threads per line
  const int groups = 256 / lpl;         // p-cells per workgroup step
  const int g = t / lpl, c0 = (t % lpl) * w;
  double a[5][VEC];
#pragma unroll
  for (int k = 0; k < 5; ++k)
#pragma unroll
    for (int c = 0; c < VEC; ++c) a[k][c] = 0.0;
  if (g < groups) {
    const T* col = f + I.off_f + c0;
    const int pf = ch.p0 + g;
    int j1 = pf % I.n_p1, j23 = pf / I.n_p1;
    int j2 = j23 % I.n_p2, j3 = j23 / I.n_p2;
    // four p-cells per pass: their loads are issued together, unconditionally
    // (an index past the chunk reads p-cell pf and is then dropped), and the
    // sums still run in increasing p
    auto pass = [&](auto vec_c) {
      constexpr bool VV = decltype(vec_c)::value;
      for (int p = pf; p < ch.p1; p += 4 * groups) {
        double x[4][VEC];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int pu = p + u * groups;
          const long row = (long)(pu < ch.p1 ? pu : pf) * n1;
          if constexpr (VV) {
            const V xv = __builtin_nontemporal_load((const V*)(col + row));
#pragma unroll
            for (int c = 0; c < VEC; ++c) x[u][c] = double(xv[c]);
          } else {
            x[u][0] = double(col[row]);
#pragma unroll
            for (int c = 1; c < VEC; ++c) x[u][c] = 0.0;
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (p + u * groups < ch.p1) {
            const double v1 = centre(I.p1s, j1, I.dp1), v2 = centre(I.p2s, j2, I.dp2);
            const double v3 = centre(I.p3s, j3, I.dp3);
            const double e = 0.5 * ((v1 * v1 + v2 * v2) + v3 * v3);
#pragma unroll
            for (int c = 0; c < VEC; ++c) {
              a[0][c] += x[u][c];
              a[1][c] += x[u][c] * v1;
              a[2][c] += x[u][c] * v2;
              a[3][c] += x[u][c] * v3;
              a[4][c] += x[u][c] * e;
            }
          }
          j1 += groups;
          while (j1 >= I.n_p1) {
            j1 -= I.n_p1;
            if (++j2 == I.n_p2) {
              j2 = 0;
              ++j3;
            }
          }
        }
      }
    };
    if (vec) pass(std::integral_constant<bool, true>());
    else pass(std::integral_constant<bool, false>());
  }
  __shared__ double cell[5][256];
  // per thread the partial of each of its cells, then a fixed-order sum over
  // the p-groups for every cell of the line
  for (int c = 0; c < w; ++c) {
#pragma unroll
    for (int k = 0; k < 5; ++k) red[k][t] = a[k][c];
    __syncthreads();
    if (t < lpl) {
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        double s = 0.0;
        for (int gg = 0; gg < groups; ++gg) s += red[k][gg * lpl + t];
        cell[k][t * w + c] = s;
      }
    }
    __syncthreads();
  }
  if (t < n1)
    for (int k = 0; k < 5; ++k) part[I.off_part + ((long)ch.chunk * 5 + k) * n1 + t] = cell[k][t];
}

// Moment block and Maxwellian tables of every instance: workgroup (instance
// blockIdx.x, table group blockIdx.y of ENS_TGROUPS), one thread per q1 cell.
// Every group sums the partials of its q-cells (fixed order); group 0 writes
// the moment block and the divergence flag; group g writes the table rows
// j = g, g + ENS_TGROUPS, ...
constexpr int ENS_TGROUPS = 8;

template <typename T>
__global__ __launch_bounds__(256) void ens_finalize_kernel(const EnsInst* inst,
                                                           const double* part, double* mom,
                                                           T* tab, int* flag) {
  const EnsInst I = inst[blockIdx.x];
  const int q = threadIdx.x;
  if (q >= I.n1) return;
  const int n1 = I.n1;
  double m[5];
  sum_moment_partials(part + I.off_part + q, I.nchunk, n1, m);
  for (int k = 0; k < 5; ++k) m[k] *= I.measure;
  double* mb = mom + I.off_mom;
  if (blockIdx.y == 0) {
    if (!isfinite(m[0])) *flag = 1;
    write_moment_block(mb, n1, q, m, I.mass, I.kB, I.p_dim);
  }
  if (!I.source) return;
  const BgkCell c = bgk_cell(m[0], m[1], m[2], m[3], m[4], I.mass, I.p_dim);
  const double pref = maxwell_prefactor(c, I.mass, I.kB, I.p_dim);
  T* tb = tab + I.off_tab;
  const int nj = I.n_p1 + I.n_p2 + I.n_p3;
  for (int j = blockIdx.y; j < nj; j += ENS_TGROUPS) {
    double val;
    if (j < I.n_p1) val = pref * maxwell_factor(centre(I.p1s, j, I.dp1), c.u1, c.T, I.mass, I.kB);
    else if (j < I.n_p1 + I.n_p2)
      val = maxwell_factor(centre(I.p2s, j - I.n_p1, I.dp2), c.u2, c.T, I.mass, I.kB);
    else val = maxwell_factor(centre(I.p3s, j - I.n_p1 - I.n_p2, I.dp3), c.u3, c.T, I.mass, I.kB);
    tb[(long)j * n1 + q] = T(val);
  }
}
#endif

}  // namespace bx
