// Ensembles of independent instances that are one-dimensional in q (periodic
// q1, N_q2 = 1, fields off), each with its own grid, velocity grid, collision
// time and time step, advanced together.  Every instance keeps f in the
// q1-line layout [P][n1] inside pooled arrays.
//
// One SSP-RK3 / RK2 stage of the whole ensemble is two launches:
//  * ens_chunk_kernel: one workgroup per (instance, p-chunk).  A chunk is a
//    contiguous span of whole q1 lines of one instance; the workgroup streams
//    it through LDS in batches of lines (16-byte loads and stores wherever
//    the rows allow, one batch's loads in flight while the previous batch's
//    moments are summed), evaluates the reconstruction, upwind flux, BGK
//    source and RK combination of every cell (df_dt_fvm.py:56-97,
//    collision_operator.py:70-128, fvm_operator.py:22-33) and sums the FP64
//    partial moments of the stage output over the chunk's lines in a fixed
//    order (compute_moments.py:7-59), so f is not read a second time;
//  * ens_finalize_kernel: one workgroup per (instance, block of q1 cells):
//    the moment block (moments.py:5-22, collision_operator.py:102-116) from
//    the partials, and the separable Maxwellian tables the next stage reads.
// ens_moments_kernel computes the same partials from a given state (the
// initial one).  The chunk list is built on the host; no kernel waits on
// another workgroup.
#pragma once
#include <type_traits>
#include "qline.h"

namespace bx {

struct EnsInst {
  int n1, P, n_p1, n_p2, n_p3, p_dim, source;
  int w;           // q1 cells per chunk-kernel item: 2 * (16 / sizeof(T)) on vector rows, else 1
  int batch;       // q1 lines a chunk workgroup holds in LDS at once
  int nchunk;      // p-chunks (partial moment blocks)
  float rn_p1, rn_p2;  // 1 / n_p1 and 1 / n_p2 for fdivmod
  long off_f;     // element offset of the [P][n1] block in the pooled f-sized arrays
  long off_tab;   // element offset of the [n_p1+n_p2+n_p3][n1] Maxwellian tables
  long off_mom;   // double offset of the [12][n1] moment block
  long off_part;  // double offset of the [nchunk][5][n1] partial moments
  double inv_dq, cdt, inv_tau, p1s, dp1, p2s, dp2, p3s, dp3, measure, mass, kB;
};

struct EnsChunk {
  int inst, chunk, p0, p1;  // p-cells p0 .. p1-1 of instance inst; partial block ``chunk``
};

struct EnsStageArgs {
  const EnsInst* inst;
  const EnsChunk* chunks;
  const void* u;
  const void* w;
  const void* un;
  void* out;
  const void* tab;
  double* part;
  double a, b, cscale;  // stage: out = a*un + b*(w + cscale*dt_i*L(u))
  int use_un, b_one;
};

// One item (a run of W q1 cells of one line) per thread and batch: a batch
// holds at most ENS_THREADS items.
constexpr int ENS_THREADS = 256;
// Periodic halo cells on each side of an LDS line: 4 keeps every line's
// cells 16-byte aligned and covers the 3-cell reach of the stencil.
constexpr int ENS_HALO = 4;

// LDS of the chunk kernel: the batch's lines of u (halo-padded), its
// outputs [batch][n1], and the moment weights (v1, v2, v3, |v|^2 / 2) of its
// lines in two slots (the previous batch's are read while the next batch's
// are written).  The group partials at the end of a chunk reuse the first
// region.
template <typename T>
__host__ __device__ inline size_t ens_chunk_lds_lines(int n1, int batch) {
  size_t fo = (size_t)batch * (2 * n1 + 2 * ENS_HALO) * sizeof(T);
  const size_t red = 5 * ENS_THREADS * sizeof(double);
  return ((fo > red ? fo : red) + 15) / 16 * 16;
}
template <typename T>
__host__ __device__ inline size_t ens_chunk_lds(int n1, int batch) {
  return ens_chunk_lds_lines<T>(n1, batch) + 2 * (size_t)batch * 4 * sizeof(double);
}

// Global and LDS accesses of N consecutive cells: 16-byte vectors when N is
// a multiple of the vector width (the caller guarantees the alignment),
// else one element.
template <typename T, int N, bool STREAM>
__device__ __forceinline__ void ens_ld(const T* p, T* x) {
  constexpr int V = 16 / sizeof(T);
  if constexpr (N % V != 0) {
    static_assert(N == 1, "one element or whole vectors");
    x[0] = STREAM ? __builtin_nontemporal_load(p) : *p;
  } else {
    typedef T Vt __attribute__((ext_vector_type(V)));
#pragma unroll
    for (int k = 0; k < N; k += V) {
      const Vt v = STREAM ? __builtin_nontemporal_load((const Vt*)(p + k)) : *(const Vt*)(p + k);
#pragma unroll
      for (int r = 0; r < V; ++r) x[k + r] = v[r];
    }
  }
}
template <typename T, int N, bool STREAM>
__device__ __forceinline__ void ens_st(T* p, const T* x) {
  constexpr int V = 16 / sizeof(T);
  if constexpr (N % V != 0) {
    static_assert(N == 1, "one element or whole vectors");
    if (STREAM) __builtin_nontemporal_store(x[0], p);
    else *p = x[0];
  } else {
    typedef T Vt __attribute__((ext_vector_type(V)));
#pragma unroll
    for (int k = 0; k < N; k += V) {
      Vt v;
#pragma unroll
      for (int r = 0; r < V; ++r) v[r] = x[k + r];
      if (STREAM) __builtin_nontemporal_store(v, (Vt*)(p + k));
      else *(Vt*)(p + k) = v;
    }
  }
}

// Velocity indices (j1, j2, j3) of p-cell p.
__device__ __forceinline__ void ens_jjj(const EnsInst& I, int p, int& j1, int& j2, int& j3) {
  const int j23 = fdivmod(p, I.n_p1, I.rn_p1, j1);
  j3 = fdivmod(j23, I.n_p2, I.rn_p2, j2);
}

// One chunk: lines ch.p0 .. ch.p1-1 of instance I in batches of I.batch
// lines; W cells per item (W = I.w).
//
// Per batch: each thread issues the loads of its item (u, w, un, table
// rows), the first nb threads put their line's moment weights in LDS, every
// thread adds the previous batch's outputs into its moment sums and stores
// its u into the halo-padded LDS lines; barrier; each item reconstructs the
// W + 1 lower faces of its cells from LDS (the indicator quantities of the
// W + 4 cells they need evaluated in the thread), forms the stage values,
// stores them to memory and to the LDS output rows; barrier.
//
// Moment order: thread t sums cell t mod n1 over the chunk lines whose
// chunk-relative index is congruent to t / n1 modulo G = 256 / n1, in
// increasing line order; the G group sums of a cell are then added in group
// order.  The order depends on n1 and the chunk only, not on the batch size.
// Stage forms (out = a*un + b*(w + cscale*dt*L(u))) the kernel is
// specialised for: w = u, a = 0, b = 1 (the first stage of either stepper);
// w = u, a != 0 (SSP-RK3 stages 2 and 3); anything else.
enum { ENS_FORM_FIRST = 0, ENS_FORM_COMBINE = 1, ENS_FORM_GENERAL = 2 };

template <typename T, int R, int W, int FORM>
__device__ void ens_chunk(const EnsStageArgs& ea, const EnsChunk& ch, const EnsInst& I,
                          char* smem) {
  const int t = threadIdx.x, n1 = I.n1, nv = n1 / W, B = I.batch;
  const int ls = n1 + 2 * ENS_HALO;
  T* F = (T*)smem;           // [B][ls], cell i of line l at l * ls + ENS_HALO + i
  T* O = F + (long)B * ls;   // [B][n1]
  double4* LW = (double4*)(smem + ens_chunk_lds_lines<T>(n1, B));  // [2][B]
  const long base = I.off_f;
  const T* u = (const T*)ea.u + base;
  const T* wv = (const T*)ea.w + base;
  const T* unv = (const T*)ea.un + base;
  T* out = (T*)ea.out + base;
  const T* tab = (const T*)ea.tab + I.off_tab;
  const bool w_is_u = FORM == ENS_FORM_GENERAL ? ea.w == ea.u : true;
  const bool use_un = FORM == ENS_FORM_GENERAL ? ea.use_un != 0 : FORM == ENS_FORM_COMBINE;
  const bool b_one = FORM == ENS_FORM_GENERAL ? ea.b_one != 0 : FORM == ENS_FORM_FIRST;
  const bool src = I.source != 0, src2 = src && I.p_dim >= 2, src3 = src && I.p_dim >= 3;
  const T tinv = T(I.inv_dq), ttau = T(I.inv_tau), tcdt = T(I.cdt * ea.cscale);
  const T ta = T(ea.a), tb = T(ea.b);
  const float rnv = 1.0f / float(nv);
  const int G = ENS_THREADS / n1;
  const int mg = t / n1, mi = t - mg * n1;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  // this thread's item: line il of the batch, cells ic .. ic + W - 1
  int ic;
  const int il = fdivmod(t, nv, rnv, ic);
  ic *= W;

  // moment sums of one batch (lines l0 .. l0+nb-1, outputs in O, weights in lw)
  auto accumulate = [&](int l0, int nb, const double4* lw) {
    if (mg >= G) return;
    const int r0 = (l0 - ch.p0) % G;
    for (int l = (mg - r0 + G) % G; l < nb; l += G) {
      const double4 v = lw[l];
      const double x = double(O[l * n1 + mi]);
      acc[0] += x;
      acc[1] += x * v.x;
      acc[2] += x * v.y;
      acc[3] += x * v.z;
      acc[4] += x * v.w;
    }
  };

  int k = 0, l0 = ch.p0;
  for (; l0 < ch.p1; l0 += B, ++k) {
    const int nb = min(B, ch.p1 - l0);
    double4* lw = LW + (k & 1) * B;
    const bool live = il < nb;
    // loaded where used: u for live items, w, un and the table rows when the
    // form and the instance read them
    T f[W], wc[W], uc[W], t1[W], t2[W], t3[W];
    int j1, j2, j3;
    ens_jjj(I, l0 + (live ? il : 0), j1, j2, j3);
    const T C = T(centre(I.p1s, j1, I.dp1));
    const long g = (long)(l0 + il) * n1 + ic;
    if (live) {
      ens_ld<T, W, true>(u + g, f);
      if (!w_is_u) ens_ld<T, W, true>(wv + g, wc);
      if (use_un) ens_ld<T, W, true>(unv + g, uc);
      if (src) ens_ld<T, W, false>(tab + (long)j1 * n1 + ic, t1);
      if (src2) ens_ld<T, W, false>(tab + (long)(I.n_p1 + j2) * n1 + ic, t2);
      if (src3) ens_ld<T, W, false>(tab + (long)(I.n_p1 + I.n_p2 + j3) * n1 + ic, t3);
    }
    if (t < nb) {
      int a1, a2, a3;
      ens_jjj(I, l0 + t, a1, a2, a3);
      const double v1 = centre(I.p1s, a1, I.dp1), v2 = centre(I.p2s, a2, I.dp2);
      const double v3 = centre(I.p3s, a3, I.dp3);
      lw[t] = make_double4(v1, v2, v3, 0.5 * ((v1 * v1 + v2 * v2) + v3 * v3));
    }
    // the previous batch (always a full one) while this batch's loads are in flight
    if (k > 0) accumulate(l0 - B, B, LW + ((k - 1) & 1) * B);
    T* row = F + il * ls + ENS_HALO;
    if (live) {
      ens_st<T, W, false>(row + ic, f);
      if constexpr (W >= ENS_HALO) {
        // periodic copies: cells 0..3 after the line, cells n1-4..n1-1 before it
        if (ic == 0) ens_st<T, ENS_HALO, false>(row + n1, f);
        if (ic == n1 - W) ens_st<T, ENS_HALO, false>(row - ENS_HALO, f + W - ENS_HALO);
      } else {
        if (ic < ENS_HALO) row[n1 + ic] = f[0];
        if (ic >= n1 - ENS_HALO) row[ic - n1] = f[0];
      }
    }
    __syncthreads();
    if (live) {
      // y[q] = u at cell ic - 4 + q of the line
      T y[W + 8];
      if constexpr (W == 1) {
#pragma unroll
        for (int q = 1; q < 8; ++q) y[q] = row[ic + q - 4];
        y[0] = y[8] = T(0);
      } else {
        ens_ld<T, ENS_HALO, false>(row + ic - ENS_HALO, y);
#pragma unroll
        for (int q = 0; q < W; ++q) y[ENS_HALO + q] = f[q];
        ens_ld<T, ENS_HALO, false>(row + ic + W, y + ENS_HALO + W);
      }
      // indicator quantities of cells ic - 2 + q, q < W + 4
      T kap[W + 4];
#pragma unroll
      for (int q = 0; q < W + 4; ++q) kap[q] = ql_kappa<R, T>(y[q + 1], y[q + 2], y[q + 3]);
      // lower face of cell ic + q (q = W: the upper face of the last cell):
      // right face of the cell below (C > 0) or left face of the cell
      T s[W + 1];
      if (C > T(0)) {
#pragma unroll
        for (int q = 0; q <= W; ++q)
          s[q] = ql_face<R, true, T>(y[q + 1], y[q + 2], y[q + 3], y[q + 4], y[q + 5], kap[q],
                                     kap[q + 1], kap[q + 2]);
      } else {
#pragma unroll
        for (int q = 0; q <= W; ++q)
          s[q] = ql_face<R, false, T>(y[q + 2], y[q + 3], y[q + 4], y[q + 5], y[q + 6],
                                      kap[q + 1], kap[q + 2], kap[q + 3]);
      }
      T o[W];
#pragma unroll
      for (int q = 0; q < W; ++q) {
        const T fq = y[q + 4];
        T d = -(C * s[q + 1] - C * s[q]) * tinv;
        if (src) {
          T fmb = t1[q];
          if (src2) fmb = fmb * t2[q];
          if (src3) fmb = fmb * t3[q];
          d = d + (fmb - fq) * ttau;
        }
        const T tv = (w_is_u ? fq : wc[q]) + d * tcdt;
        o[q] = use_un ? ta * uc[q] + tb * tv : (b_one ? tv : tb * tv);
      }
      ens_st<T, W, true>(out + g, o);
      ens_st<T, W, false>(O + il * n1 + ic, o);
    }
    __syncthreads();
  }
  // the last batch, then the group sums in group order
  accumulate(l0 - B, ch.p1 - (l0 - B), LW + ((k - 1) & 1) * B);
  __syncthreads();
  double* red = (double*)smem;  // [G][5][n1]
  if (mg < G)
#pragma unroll
    for (int q = 0; q < 5; ++q) red[(mg * 5 + q) * n1 + mi] = acc[q];
  __syncthreads();
  if (t < n1) {
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      double sum = 0.0;
      for (int g2 = 0; g2 < G; ++g2) sum += red[(g2 * 5 + q) * n1 + t];
      ea.part[I.off_part + ((long)ch.chunk * 5 + q) * n1 + t] = sum;
    }
  }
}

#ifndef BX_NO_KERNELS
template <typename T, int R, int FORM>
__global__ __launch_bounds__(ENS_THREADS) void ens_chunk_kernel(EnsStageArgs ea) {
  extern __shared__ __align__(16) char ens_smem[];
  const EnsChunk ch = ea.chunks[blockIdx.x];
  const EnsInst I = ea.inst[ch.inst];
  if (I.w > 1) ens_chunk<T, R, 32 / sizeof(T), FORM>(ea, ch, I, ens_smem);
  else ens_chunk<T, R, 1, FORM>(ea, ch, I, ens_smem);
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

// Moment block and Maxwellian tables of every instance: one workgroup per
// (instance, block of ENS_FIN_CELLS q1 cells), so every partial is read
// once.  Thread t works on cell c0 + (t mod ENS_FIN_CELLS) with sub-index
// s = t / ENS_FIN_CELLS < ENS_FIN_SUBS: it sums the partials of chunks
// s, s + ENS_FIN_SUBS, ... in increasing order, and the sub-sums of a cell
// are added in sub-index order (an order fixed by the chunk count).  The
// first sub writes the cell's moment block and the divergence flag; every
// sub then writes the table rows j = s, s + ENS_FIN_SUBS, ... of the cell.
constexpr int ENS_FIN_CELLS = 16;
constexpr int ENS_FIN_SUBS = 16;

template <typename T>
__global__ __launch_bounds__(ENS_FIN_CELLS * ENS_FIN_SUBS) void ens_finalize_kernel(
    const EnsInst* inst, const int2* blocks, const double* part, double* mom, T* tab, int* flag) {
  __shared__ double red[5][ENS_FIN_SUBS][ENS_FIN_CELLS];
  __shared__ double cell[5][ENS_FIN_CELLS];  // prefactor, u1, u2, u3, T
  const int2 blk = blocks[blockIdx.x];       // (instance, first cell)
  const EnsInst I = inst[blk.x];
  const int n1 = I.n1, t = threadIdx.x;
  const int sub = t / ENS_FIN_CELLS, ci = t - sub * ENS_FIN_CELLS, q = blk.y + ci;
  const bool on = q < n1;
  double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (on) {
    const double* pp = part + I.off_part + q;
    const long step = (long)ENS_FIN_SUBS * 5 * n1;
    int c = sub;
    // two chunks' loads in flight at a time, summed in chunk order
    for (; c + ENS_FIN_SUBS < I.nchunk; c += 2 * ENS_FIN_SUBS) {
      double v[2][5];
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int k = 0; k < 5; ++k) v[j][k] = pp[(long)c * 5 * n1 + j * step + (long)k * n1];
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int k = 0; k < 5; ++k) a[k] += v[j][k];
    }
    if (c < I.nchunk)
#pragma unroll
      for (int k = 0; k < 5; ++k) a[k] += pp[((long)c * 5 + k) * n1];
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) red[k][sub][ci] = a[k];
  __syncthreads();
  if (sub == 0 && on) {
    double m[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      double sum = 0.0;
#pragma unroll
      for (int s2 = 0; s2 < ENS_FIN_SUBS; ++s2) sum += red[k][s2][ci];
      m[k] = sum * I.measure;
    }
    if (!isfinite(m[0])) *flag = 1;
    write_moment_block(mom + I.off_mom, n1, q, m, I.mass, I.kB, I.p_dim);
    const BgkCell bc = bgk_cell(m[0], m[1], m[2], m[3], m[4], I.mass, I.p_dim);
    cell[0][ci] = maxwell_prefactor(bc, I.mass, I.kB, I.p_dim);
    cell[1][ci] = bc.u1;
    cell[2][ci] = bc.u2;
    cell[3][ci] = bc.u3;
    cell[4][ci] = bc.T;
  }
  __syncthreads();
  if (!I.source || !on) return;
  const double pref = cell[0][ci], u1 = cell[1][ci], u2 = cell[2][ci], u3 = cell[3][ci];
  const double Tc = cell[4][ci];
  T* tb = tab + I.off_tab;
  const int nj = I.n_p1 + I.n_p2 + I.n_p3;
  for (int j = sub; j < nj; j += ENS_FIN_SUBS) {
    double val;
    if (j < I.n_p1) val = pref * maxwell_factor(centre(I.p1s, j, I.dp1), u1, Tc, I.mass, I.kB);
    else if (j < I.n_p1 + I.n_p2)
      val = maxwell_factor(centre(I.p2s, j - I.n_p1, I.dp2), u2, Tc, I.mass, I.kB);
    else val = maxwell_factor(centre(I.p3s, j - I.n_p1 - I.n_p2, I.dp3), u3, Tc, I.mass, I.kB);
    tb[(long)j * n1 + q] = T(val);
  }
}
#endif

}  // namespace bx
