// gfx950 kernels and C-ABI of the nonlinear FVM step (see include/boltx.h).
//
// Work decomposition: one workgroup per q-cell (all P of it), so the velocity
// moments of a stage's output are a deterministic in-workgroup reduction and
// the p-direction stencils (fields on) read the q-cell's block from LDS.  The
// q-direction stencil reads neighbouring q-cells' p-rows straight from HBM/L2;
// workgroup ids are remapped so each XCD sweeps one contiguous band of q2 rows
// and its L2 holds the rows the stencil revisits.
#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>

#include "../../include/boltx.h"
#include "fvm_stage.h"
#include "qcell.h"
#include "qsweep.h"
#include "ql_launch.h"
#include "qline_moments.h"
#include "adjoint.h"
#include "adjoint_lines.h"
#include "bx_status.h"

namespace bx {

static thread_local char g_err[512];

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

// blockIdx -> q-cell: consecutive workgroups are dispatched round-robin over
// the 8 XCDs; give XCD x the contiguous range [x*Q/8, (x+1)*Q/8).
__device__ inline int xcd_map(int b, int nb) {
  if (nb % 8 != 0) return b;
  return (b % 8) * (nb / 8) + b / 8;
}

// Sum of ``v`` over the workgroup, deterministic (butterfly within waves,
// then a fixed-order sum over waves).  Result valid in thread 0.
template <int N>
__device__ inline void block_sum(double (&v)[N], double (*red)[kWaves]) {
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double x = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    v[k] = x;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < N; ++k) red[k][wave] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
      double s = red[k][0];
      for (int w = 1; w < kWaves; ++w) s += red[k][w];
      v[k] = s;
    }
  }
}

// Maxwellian factor tables of q-cell q into LDS (collision_operator.py:52-65).
template <typename T>
__device__ inline void build_tables(const StageParams& sp, int q, T* tab, QCell<T>& c) {
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

// Accumulate the five built-in moments of value x at velocity v
// (moments.py:5-22): f, f v1, f v2, f v3, 0.5 f (v1^2 + v2^2 + v3^2).
__device__ inline void add_moments(double (&acc)[5], double x, const double v[3]) {
  acc[0] += x;
  acc[1] += x * v[0];
  acc[2] += x * v[1];
  acc[3] += x * v[2];
  acc[4] += (0.5 * x) * ((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
}

template <typename T, int RQ, int RP, bool PADV>
__global__ __launch_bounds__(kBlock) void fvm_stage_kernel(StageParams sp) {
  extern __shared__ unsigned char smem[];
  __shared__ double red[5][kWaves];
  T* tab = reinterpret_cast<T*>(smem);
  T* blk = tab + (sp.n_p1 + sp.n_p2 + sp.n_p3);
  const int q = xcd_map(blockIdx.x, gridDim.x);
  QCell<T> c;
  qcell_neighbours(sp, q, c);
  c.pref = 0.0;
  c.e1 = c.e2 = c.e3 = tab;
  if (sp.source) build_tables(sp, q, tab, c);
  if constexpr (PADV) {
    if (!sp.cp1.ptr) qcell_fields(sp, c);
    const T* u = (const T*)sp.u + (long)q * sp.P;
    for (int p = threadIdx.x; p < sp.P; p += blockDim.x) blk[p] = u[p];
  }
  __syncthreads();
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  T* out = (T*)sp.out + (long)q * sp.P;
  for (int p = threadIdx.x; p < sp.P; p += blockDim.x) {
    double v[3];
    const T x = stage_cell<T, RQ, RP, PADV>(sp, c, p, blk, v);
    out[p] = x;
    add_moments(acc, double(x), v);
  }
  if (sp.mom_out) {
    block_sum(acc, red);
    if (threadIdx.x == 0) {
      for (int k = 0; k < 5; ++k) acc[k] *= sp.measure;
      write_moment_block(sp.mom_out, sp.Q, q, acc, sp.mass, sp.kB, sp.p_dim);
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void moments_kernel(StageParams sp) {
  __shared__ double red[5][kWaves];
  const int q = xcd_map(blockIdx.x, gridDim.x);
  const T* f = (const T*)sp.u + (long)q * sp.P;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int p = threadIdx.x; p < sp.P; p += blockDim.x) {
    const int j1 = p % sp.n_p1, j23 = p / sp.n_p1;
    const double v[3] = {centre(sp.p1s, j1, sp.dp1), centre(sp.p2s, j23 % sp.n_p2, sp.dp2),
                         centre(sp.p3s, j23 / sp.n_p2, sp.dp3)};
    add_moments(acc, double(f[p]), v);
  }
  block_sum(acc, red);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 5; ++k) acc[k] *= sp.measure;
    write_moment_block(sp.mom_out, sp.Q, q, acc, sp.mass, sp.kB, sp.p_dim);
  }
}

// Slab order of the p-cells for qcell_kernel: grouped by the signs of
// (v1, v2) -- (<=0,<=0), (>0,<=0), (<=0,>0), (>0,>0) -- each group in natural
// order and padded with -1 to a multiple of 64; jpack holds j1 | j2<<10 | j3<<20
// of each entry (0 for padding).  One wave; 64 p-cells per step are tested
// together and the members of the group are packed in order by a ballot.
// Run once per problem.
__global__ __launch_bounds__(64) void slab_perm_kernel(int* perm, int* jpack, int n_p1, int n_p2,
                                                       int n_p3, double p1s, double dp1, double p2s,
                                                       double dp2) {
  if (blockIdx.x != 0) return;
  const int lane = threadIdx.x;
  const int P = n_p1 * n_p2 * n_p3;
  int k = 0;  // entries written so far (uniform across the wave)
  for (int grp = 0; grp < 4; ++grp) {
    const bool pos1 = grp & 1, pos2 = grp >> 1;
    const int start = k;
    for (int base = 0; base < P; base += 64) {
      const int p = base + lane;
      int j1 = 0, j2 = 0, j3 = 0;
      bool in = false;
      if (p < P) {
        j1 = p % n_p1;
        j2 = (p / n_p1) % n_p2;
        j3 = p / (n_p1 * n_p2);
        in = (centre(p1s, j1, dp1) > 0.0) == pos1 && (centre(p2s, j2, dp2) > 0.0) == pos2;
      }
      const unsigned long long mask = __ballot(in);
      if (in) {
        const int at = k + __popcll(mask & ((1ull << lane) - 1ull));
        perm[at] = p;
        jpack[at] = j1 | (j2 << 10) | (j3 << 20);
      }
      k += __popcll(mask);
    }
    if (k > start) {
      const int pad = (64 - k % 64) % 64;
      if (lane < pad) {
        perm[k + lane] = -1;
        jpack[k + lane] = 0;
      }
      k += pad;
    }
  }
}

// J1 = e * sum(f_left v1) dV, J2 = e * sum(f_bot v2) dV, J3 = e * sum(f v3) dV
// with f_left / f_bot the upwind states of the q1 / q2 left faces
// (df_dt_fvm.py:79-80, 214-224).
// Upwind state of the lower face of the centre of a[k] = f[i-3+k]
// (riemann.py:46-49): the right face of cell i-1 when C > 0, else the left
// face of cell i; one reconstruction from the window selected per lane.
template <int RQ, typename T>
__device__ inline T lower_face_state(const T a[7], bool pos) {
  T w[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) w[k] = pos ? a[k] : a[k + 1];
  return face_value<RQ>(w[0], w[1], w[2], w[3], w[4], pos);
}

template <typename T, int RQ>
__global__ __launch_bounds__(kBlock) void currents_kernel(StageParams sp, double* J) {
  __shared__ double red[3][kWaves];
  const int q = xcd_map(blockIdx.x, gridDim.x);
  QCell<T> c;
  qcell_neighbours(sp, q, c);
  const T* u = (const T*)sp.u;
  const long P = sp.P;
  constexpr int r = recon_reach(RQ) + 1;
  double acc[3] = {0.0, 0.0, 0.0};
  const float rcp1 = 1.0f / float(sp.n_p1), rcp2 = 1.0f / float(sp.n_p2);
  for (int p = threadIdx.x; p < sp.P; p += blockDim.x) {
    int j1, j23, j2;
    j23 = qs_divmod(p, sp.n_p1, rcp1, j1);
    const int j3 = qs_divmod(j23, sp.n_p2, rcp2, j2);
    const double v1 = centre(sp.p1s, j1, sp.dp1);
    const double v2 = centre(sp.p2s, j2, sp.dp2);
    const double v3 = centre(sp.p3s, j3, sp.dp3);
    const T f = u[(long)q * P + p];
    T s1 = f, s2 = f;
    if (sp.n_q1 > 1) {
      T a[7];
      for (int k = 0; k < 7; ++k) a[k] = f;
      for (int k = 3 - r; k <= 3 + r; ++k) a[k] = k == 3 ? f : u[(long)c.nb1[k] * P + p];
      s1 = lower_face_state<RQ, T>(a, T(v1) > T(0));
    }
    if (sp.n_q2 > 1) {
      T a[7];
      for (int k = 0; k < 7; ++k) a[k] = f;
      for (int k = 3 - r; k <= 3 + r; ++k) a[k] = k == 3 ? f : u[(long)c.nb2[k] * P + p];
      s2 = lower_face_state<RQ, T>(a, T(v2) > T(0));
    }
    acc[0] += double(s1) * v1;
    acc[1] += double(s2) * v2;
    acc[2] += double(f) * v3;
  }
  block_sum(acc, red);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 3; ++k) J[k * sp.Q + q] = sp.charge * (acc[k] * sp.measure);
  }
}

// FDTD on the periodic Yee grid (fdtd_explicit.py:9-138).  Component k of
// yee / cc / J is the [Q] slice k*Q.., index i1 + n_q1*i2.
struct Grid2 {
  int n1, n2;
  __device__ int at(int i1, int i2) const {
    i1 = (i1 + n1) % n1;
    i2 = (i2 + n2) % n2;
    return i1 + n1 * i2;
  }
};

__global__ void fdtd_e_kernel(Grid2 g, double* y, const double* J, double dq1, double dq2,
                              double dt, double eps, double mu) {
  const int Q = g.n1 * g.n2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Q) return;
  const int i1 = t % g.n1, i2 = t / g.n1;
  const double* B1 = y + 3 * Q;
  const double* B2 = y + 4 * Q;
  const double* B3 = y + 5 * Q;
  const double curlB_1 = (B3[g.at(i1, i2 + 1)] - B3[t]) / dq2;
  const double curlB_2 = -(B3[g.at(i1 + 1, i2)] - B3[t]) / dq1;
  const double curlB_3 = (B2[g.at(i1 + 1, i2)] - B2[t]) / dq1 - (B1[g.at(i1, i2 + 1)] - B1[t]) / dq2;
  const double c = dt / (mu * eps);
  y[t] = y[t] + (c * curlB_1 - J[t] * dt / eps);
  y[Q + t] = y[Q + t] + (c * curlB_2 - J[Q + t] * dt / eps);
  y[2 * Q + t] = y[2 * Q + t] + (c * curlB_3 - J[2 * Q + t] * dt / eps);
}

__global__ void fdtd_b_kernel(Grid2 g, double* y, double dq1, double dq2, double dt) {
  const int Q = g.n1 * g.n2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Q) return;
  const int i1 = t % g.n1, i2 = t / g.n1;
  const double* E1 = y;
  const double* E2 = y + Q;
  const double* E3 = y + 2 * Q;
  const double curlE_1 = (E3[t] - E3[g.at(i1, i2 - 1)]) / dq2;
  const double curlE_2 = -(E3[t] - E3[g.at(i1 - 1, i2)]) / dq1;
  const double curlE_3 = (E2[t] - E2[g.at(i1 - 1, i2)]) / dq1 - (E1[t] - E1[g.at(i1, i2 - 1)]) / dq2;
  y[3 * Q + t] = y[3 * Q + t] + -dt * curlE_1;
  y[4 * Q + t] = y[4 * Q + t] + -dt * curlE_2;
  y[5 * Q + t] = y[5 * Q + t] + -dt * curlE_3;
}

// fields.py:527-557.
__global__ void yee_to_cc_kernel(Grid2 g, const double* y, double* cc) {
  const int Q = g.n1 * g.n2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Q) return;
  const int i1 = t % g.n1, i2 = t / g.n1;
  const int p1 = g.at(i1 + 1, i2), p2 = g.at(i1, i2 + 1), p12 = g.at(i1 + 1, i2 + 1);
  cc[t] = 0.5 * (y[t] + y[p1]);
  cc[Q + t] = 0.5 * (y[Q + t] + y[Q + p2]);
  cc[2 * Q + t] = y[2 * Q + t];
  cc[3 * Q + t] = 0.5 * (y[3 * Q + t] + y[3 * Q + p2]);
  cc[4 * Q + t] = 0.5 * (y[4 * Q + t] + y[4 * Q + p1]);
  cc[5 * Q + t] = 0.25 * (y[5 * Q + t] + y[5 * Q + p2] + y[5 * Q + p1] + y[5 * Q + p12]);
}

__global__ void check_finite_kernel(const double* n, int Q, int* flag) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < Q && !isfinite(n[t])) flag[0] = 1;
}

// Ghost q-rows of the non-periodic sides (boundaries.py:147-311, 415-577 after
// the zero fill of a ghosted DMDA axis): one thread per (ghost row, p).
struct GhostParams {
  int n_q1, n_q2, n_p1, n_p2, P, Q, ng, gh1, rows;
  int bc[4];
  double p1s, dp1, p2s, dp2;
  const void* dir[4];
};

template <typename T>
__global__ void fill_ghosts_kernel(GhostParams g, T* f) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.rows * g.P) return;
  const int r = (int)(t / g.P), p = (int)(t % g.P);
  // side, ghost cell c and the cell index along the boundary
  int side, c, io, local;
  const int b1 = g.gh1 ? 2 * g.ng * g.n_q2 : 0;
  if (r < b1) {
    side = r < g.ng * g.n_q2 ? 0 : 1;
    local = r - side * g.ng * g.n_q2;
    c = local / g.n_q2;
    io = local % g.n_q2;
  } else {
    const int rr = r - b1;
    side = rr < g.ng * g.n_q1 ? 2 : 3;
    local = rr - (side - 2) * g.ng * g.n_q1;
    c = local / g.n_q1;
    io = local % g.n_q1;
  }
  const int kind = g.bc[side];
  const int j1 = p % g.n_p1, j2 = (p / g.n_p1) % g.n_p2;
  T v = T(0);  // a ghosted DMDA boundary zone is never written: zero
  if (kind == BOLTX_BC_MIRROR || kind == BOLTX_BC_MIRROR_DIRICHLET) {
    // mirror image of interior cell ng-1-c (left/bottom) or n-1-c (right/top),
    // normal velocity reversed
    int qs, ps;
    if (side < 2) {
      const int i1 = side == 0 ? g.ng - 1 - c : g.n_q1 - 1 - c;
      qs = i1 + g.n_q1 * io;
      ps = p + (g.n_p1 - 1 - 2 * j1);
    } else {
      const int i2 = side == 2 ? g.ng - 1 - c : g.n_q2 - 1 - c;
      qs = io + g.n_q1 * i2;
      ps = p + (g.n_p2 - 1 - 2 * j2) * g.n_p1;
    }
    v = f[(long)qs * g.P + ps];
  }
  if (kind == BOLTX_BC_DIRICHLET || kind == BOLTX_BC_MIRROR_DIRICHLET) {
    // inflow: C_q1 = p1 > 0 on the left, < 0 on the right; C_q2 = p2 likewise
    const double C = side < 2 ? centre(g.p1s, j1, g.dp1) : centre(g.p2s, j2, g.dp2);
    const bool in = (side % 2 == 0) ? C > 0.0 : C < 0.0;
    if (in) v = ((const T*)g.dir[side])[(long)local * g.P + p];
  }
  f[(long)(g.Q + r) * g.P + p] = v;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
int check_grid(const boltx_grid* g) {
  if (!g) return fail(BOLTX_EINVAL, "grid is NULL");
  if (g->n_q1 < 1 || g->n_q2 < 1 || g->n_p1 < 1 || g->n_p2 < 1 || g->n_p3 < 1)
    return fail(BOLTX_EINVAL, "grid sizes must be positive (got q %d x %d, p %d x %d x %d)",
                g->n_q1, g->n_q2, g->n_p1, g->n_p2, g->n_p3);
  if (g->n_species != 1)
    return fail(BOLTX_EUNSUPPORTED, "n_species = %d: only one species is supported", g->n_species);
  if (g->layout != BOLTX_LAYOUT_REFERENCE && g->layout != BOLTX_LAYOUT_Q1_LINES)
    return fail(BOLTX_EINVAL, "layout = %d", g->layout);
  if (g->layout == BOLTX_LAYOUT_Q1_LINES && g->n_q1 > 256)
    return fail(BOLTX_EUNSUPPORTED, "the q1-line layout holds at most 256 q1 cells (got %d)",
                g->n_q1);
  for (int k = 0; k < 4; ++k)
    if (g->bc[k] < BOLTX_BC_PERIODIC || g->bc[k] > BOLTX_BC_MIRROR_DIRICHLET)
      return fail(BOLTX_EINVAL, "bc[%d] = %d", k, g->bc[k]);
  for (int a = 0; a < 2; ++a)
    if ((g->bc[2 * a] == BOLTX_BC_PERIODIC) != (g->bc[2 * a + 1] == BOLTX_BC_PERIODIC))
      return fail(BOLTX_EINVAL, "a periodic q axis is periodic on both sides");
  const bool gh1 = g->bc[0] != BOLTX_BC_PERIODIC, gh2 = g->bc[2] != BOLTX_BC_PERIODIC;
  if (gh1 || gh2) {
    if (g->layout != BOLTX_LAYOUT_REFERENCE)
      return fail(BOLTX_EUNSUPPORTED, "non-periodic q boundaries need the reference layout");
    if (g->n_ghost < 1 || g->n_ghost > 3)
      return fail(BOLTX_EINVAL, "n_ghost = %d (1..3 with non-periodic boundaries)", g->n_ghost);
    if ((gh1 && g->n_q1 < g->n_ghost) || (gh2 && g->n_q2 < g->n_ghost))
      return fail(BOLTX_EINVAL, "a non-periodic q axis needs at least n_ghost cells");
  }
  const long P = (long)g->n_p1 * g->n_p2 * g->n_p3;
  const long Q = (long)g->n_q1 * g->n_q2;
  if (P * Q > (1L << 40)) return fail(BOLTX_EINVAL, "grid too large");
  // p-indices are split into (j1, j2, j3) through FP32 reciprocals
  if (P >= (1L << 24)) return fail(BOLTX_EINVAL, "2^24 or more p-cells per q-cell");
  if (Q > 0x7fffffff) return fail(BOLTX_EINVAL, "too many q-cells");
  return BOLTX_OK;
}

// Ghost q-rows of the non-periodic axes (0 when q is periodic).
static long ghost_rows(const boltx_grid* g) {
  long n = 0;
  if (g->bc[0] != BOLTX_BC_PERIODIC) n += 2L * g->n_ghost * g->n_q2;
  if (g->bc[2] != BOLTX_BC_PERIODIC) n += 2L * g->n_ghost * g->n_q1;
  return n;
}

static StageParams base_params(const boltx_grid* g) {
  StageParams sp;
  memset(&sp, 0, sizeof(sp));
  sp.n_q1 = g->n_q1;
  sp.n_q2 = g->n_q2;
  sp.n_p1 = g->n_p1;
  sp.n_p2 = g->n_p2;
  sp.n_p3 = g->n_p3;
  sp.P = g->n_p1 * g->n_p2 * g->n_p3;
  sp.Q = g->n_q1 * g->n_q2;
  sp.p1s = g->p1_start;
  sp.p2s = g->p2_start;
  sp.p3s = g->p3_start;
  sp.dp1 = g->dp1;
  sp.dp2 = g->dp2;
  sp.dp3 = g->dp3;
  sp.inv_dq1 = 1.0 / g->dq1;
  sp.inv_dq2 = 1.0 / g->dq2;
  sp.inv_dp1 = 1.0 / g->dp1;
  sp.inv_dp2 = 1.0 / g->dp2;
  sp.inv_dp3 = 1.0 / g->dp3;
  sp.measure = g->dp3 * g->dp2 * g->dp1;
  sp.gh1 = g->bc[0] != BOLTX_BC_PERIODIC;
  sp.gh2 = g->bc[2] != BOLTX_BC_PERIODIC;
  sp.ng = (sp.gh1 || sp.gh2) ? g->n_ghost : 0;
  return sp;
}

int launch_status(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(BOLTX_ELAUNCH, "%s: %s", what, hipGetErrorString(e));
  return BOLTX_OK;
}

template <typename T, int RQ, int RP, bool PADV>
static int launch_stage_t(const StageParams& sp, hipStream_t s) {
  size_t lds = sizeof(T) * (size_t)(sp.n_p1 + sp.n_p2 + sp.n_p3);
  if (PADV) lds += sizeof(T) * (size_t)sp.P;
  if (lds > 160 * 1024) return fail(BOLTX_EUNSUPPORTED, "p-block of %zu bytes exceeds LDS", lds);
  hipLaunchKernelGGL((fvm_stage_kernel<T, RQ, RP, PADV>), dim3(sp.Q), dim3(kBlock), lds, s, sp);
  return launch_status("fvm_stage_kernel");
}

template <typename T, int RQ>
static int launch_stage_rp(const StageParams& sp, int rp, bool padv, hipStream_t s) {
  if (!padv) return launch_stage_t<T, RQ, RC_PC, false>(sp, s);
  switch (rp) {
    case RC_PC: return launch_stage_t<T, RQ, RC_PC, true>(sp, s);
    case RC_MINMOD: return launch_stage_t<T, RQ, RC_MINMOD, true>(sp, s);
    case RC_PPM: return launch_stage_t<T, RQ, RC_PPM, true>(sp, s);
    case RC_WENO5: return launch_stage_t<T, RQ, RC_WENO5, true>(sp, s);
  }
  return fail(BOLTX_EINVAL, "recon_p = %d", rp);
}

template <typename T>
static int launch_stage_rq(const StageParams& sp, int rq, int rp, bool padv, hipStream_t s) {
  switch (rq) {
    case RC_PC: return launch_stage_rp<T, RC_PC>(sp, rp, padv, s);
    case RC_MINMOD: return launch_stage_rp<T, RC_MINMOD>(sp, rp, padv, s);
    case RC_PPM: return launch_stage_rp<T, RC_PPM>(sp, rp, padv, s);
    case RC_WENO5: return launch_stage_rp<T, RC_WENO5>(sp, rp, padv, s);
  }
  return fail(BOLTX_EINVAL, "recon_q = %d", rq);
}

// Slab count of the sign-grouped p order (must agree with slab_perm_kernel).
static int qc_nslab(const boltx_grid* g) {
  int n1neg = 0, n2neg = 0;
  for (int i = 0; i < g->n_p1; ++i) n1neg += centre(g->p1_start, i, g->dp1) > 0.0 ? 0 : 1;
  for (int i = 0; i < g->n_p2; ++i) n2neg += centre(g->p2_start, i, g->dp2) > 0.0 ? 0 : 1;
  const long n1[2] = {n1neg, g->n_p1 - n1neg}, n2[2] = {n2neg, g->n_p2 - n2neg};
  long slabs = 0;
  for (int grp = 0; grp < 4; ++grp) slabs += (n1[grp & 1] * n2[grp >> 1] * g->n_p3 + 63) / 64;
  return (int)slabs;
}

// The slab-ordered kernel needs per-axis p sizes below 1024 (packed indices).
static bool qc_applies(const boltx_grid* g) {
  return g->n_p1 < 1024 && g->n_p2 < 1024 && g->n_p3 < 1024;
}

static size_t qc_workspace_bytes(int nslab) { return 2 * (size_t)nslab * 64 * sizeof(int); }

// LDS bytes of qsweep_kernel: tables, the padded p-block and its accumulator.
static size_t qsweep_lds(const StageParams& sp) {
  const size_t pp = (size_t)(sp.n_p1 + 1) * sp.n_p2 * sp.n_p3;
  return sizeof(float) * ((size_t)(sp.n_p1 + sp.n_p2 + sp.n_p3) + 2 * pp);
}

template <typename T, int RQ, int RP, bool PADV>
static int launch_qc_t(const QCParams& qp, hipStream_t s) {
  const StageParams& sp = qp.sp;
  if constexpr (sizeof(T) == 4 && PADV && RP == RC_WENO5) {
    // FP32 WENO5 p-advection: the p-line sweeps, when the padded block and
    // the accumulator fit one CU's LDS (one 16-wave workgroup per CU)
    const size_t lq = qsweep_lds(sp);
    if (lq + 1024 <= 160 * 1024) {
      hipLaunchKernelGGL((qsweep_kernel<RQ, 1024>), dim3(sp.Q), dim3(1024), lq, s, qp);
      return launch_status("qsweep_kernel");
    }
  }
  size_t lds = sizeof(T) * (size_t)(sp.n_p1 + sp.n_p2 + sp.n_p3);
  if (PADV) lds += sizeof(T) * (size_t)sp.P;
  if (lds > 160 * 1024) return fail(BOLTX_EUNSUPPORTED, "p-block of %zu bytes exceeds LDS", lds);
  if constexpr (!PADV) {
    hipLaunchKernelGGL((qcell_kernel<T, RQ, RP, false, 256>), dim3(sp.Q), dim3(256), lds, s, qp);
  } else {
    if (2 * (lds + 1024) <= 160 * 1024) {
      // two 8-wave workgroups per CU (the 1024 B cover the kernel's own LDS)
      hipLaunchKernelGGL((qcell_kernel<T, RQ, RP, true, 512>), dim3(sp.Q), dim3(512), lds, s, qp);
    } else {
      hipLaunchKernelGGL((qcell_kernel<T, RQ, RP, true, 1024>), dim3(sp.Q), dim3(1024), lds, s, qp);
    }
  }
  return launch_status("qcell_kernel");
}

template <typename T, int RQ>
static int launch_qc_rp(const QCParams& qp, int rp, bool padv, hipStream_t s) {
  if (!padv) return launch_qc_t<T, RQ, RC_PC, false>(qp, s);
  switch (rp) {
    case RC_PC: return launch_qc_t<T, RQ, RC_PC, true>(qp, s);
    case RC_MINMOD: return launch_qc_t<T, RQ, RC_MINMOD, true>(qp, s);
    case RC_PPM: return launch_qc_t<T, RQ, RC_PPM, true>(qp, s);
    default: return launch_qc_t<T, RQ, RC_WENO5, true>(qp, s);
  }
}

template <typename T>
static int launch_qc(const QCParams& qp, int rq, int rp, bool padv, hipStream_t s) {
  switch (rq) {
    case RC_PC: return launch_qc_rp<T, RC_PC>(qp, rp, padv, s);
    case RC_MINMOD: return launch_qc_rp<T, RC_MINMOD>(qp, rp, padv, s);
    case RC_PPM: return launch_qc_rp<T, RC_PPM>(qp, rp, padv, s);
    default: return launch_qc_rp<T, RC_WENO5>(qp, rp, padv, s);
  }
}

// ---- q1-line layout --------------------------------------------------------
constexpr int ML_CHUNKS = 8;

// Number of q2 segments for ``groups`` workgroups per segment when ``slots``
// workgroups are resident at once: segments of about 32 rows (a segment
// re-reads a 4-row window halo, so longer ones read less), split in two
// while the launch would not fill the resident slots once, down to 4 rows.
// Measured at 128^2 q x 48^2 p on an MI355X (one-step graph replays): 4
// segments 0.740 ms per FP64 step, 8 segments 0.752, 2 segments 0.782, 6
// (uneven) 0.764; FP32-bulk 0.383-0.386 for 4 and 8.  BOLTX_QL_SEGMENTS
// overrides the count (tuning sweeps).
static int ql_segments(int n2, long groups, long slots) {
  if (const char* e = getenv("BOLTX_QL_SEGMENTS")) {
    const int v = atoi(e);
    if (v >= 1) return v < n2 ? v : n2;
  }
  if (n2 <= 1) return 1;
  int ns = n2 / 32 > 1 ? n2 / 32 : 1;
  while (groups * ns < slots && n2 / (2 * ns) >= 4) ns *= 2;
  return ns;
}

// Resident q1-line workgroups on the current device for the kernel picked
// by (precision, n1, recon, mode, pd), cached per device and kernel.
template <typename T>
static long ql_slots(int n1, int recon, int mode, int pd) {
  static std::mutex mu;
  static std::map<std::array<int, 5>, long> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  const int nrc = n1 <= 64 ? 1 : n1 <= 128 ? 2 : 4;
  const std::array<int, 5> key{dev, nrc, recon, mode, pd};
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 0;
  const long v = (long)cus * ql_blocks_per_cu<T>(n1, recon, mode, pd);
  cache[key] = v;
  return v;
}

static size_t ql_table_bytes(const boltx_grid* g, int precision) {
  const size_t nj = (size_t)g->n_p1 + g->n_p2 + g->n_p3;
  const size_t Q = (size_t)g->n_q1 * g->n_q2;
  return (nj * Q * (precision == BOLTX_FP64 ? 8 : 4) + 255) / 256 * 256;
}

static size_t ql_workspace_bytes(const boltx_grid* g, int precision) {
  return ql_table_bytes(g, precision) + (size_t)ML_CHUNKS * 5 * g->n_q1 * g->n_q2 * sizeof(double);
}

template <typename T>
static int ql_moments(const boltx_grid* g, const StageParams& sp, const void* f, double* mom,
                      double* part, hipStream_t s) {
  if (mline_vec_applies<T>(g->n_q1)) {
    hipLaunchKernelGGL(mline_vec_kernel<T>, dim3(g->n_q2 * ML_CHUNKS), dim3(256), 0, s,
                       (const T*)f, g->n_q1, g->n_q2, sp.P, ML_CHUNKS, g->n_p1, g->n_p2,
                       g->p1_start, g->dp1, g->p2_start, g->dp2, g->p3_start, g->dp3, part);
  } else {
    const int nblk = (g->n_q1 + 63) / 64;
    hipLaunchKernelGGL(mline_kernel<T>, dim3(nblk * g->n_q2 * ML_CHUNKS), dim3(256), 0, s,
                       (const T*)f, g->n_q1, g->n_q2, sp.P, ML_CHUNKS, g->n_p1, g->n_p2,
                       g->p1_start, g->dp1, g->p2_start, g->dp2, g->p3_start, g->dp3, part);
  }
  int rc = launch_status("mline_kernel");
  if (rc) return rc;
  // one wave per workgroup: Q threads spread over every CU
  hipLaunchKernelGGL(ml_finalize_kernel, dim3((sp.Q + 63) / 64), dim3(64), 0, s, part,
                     ML_CHUNKS, sp.Q, sp.measure, mom, sp.mass, sp.kB, sp.p_dim);
  return launch_status("ml_finalize_kernel");
}

template <typename T>
static int ql_stage(const boltx_grid* g, const boltx_physics* ph, const boltx_stage* st,
                    void* workspace, const StageParams& sp, hipStream_t s) {
  if (g->n_q1 > 1 && g->n_q1 < QL_HALO)
    return fail(BOLTX_EUNSUPPORTED, "the q1-line layout needs n_q1 = 1 or n_q1 >= %d", QL_HALO);
  char* ws = (char*)workspace;
  T* tab = (T*)ws;
  double* part = (double*)(ws + ql_table_bytes(g, ph->precision));
  int rc;
  if (sp.source) {
    const dim3 tg((unsigned)((sp.Q + 255) / 256), (unsigned)(g->n_p1 + g->n_p2 + g->n_p3));
    hipLaunchKernelGGL(bgk_table_kernel<T>, tg, dim3(256), 0, s,
                       st->mom_in, sp.Q, g->n_p1, g->n_p2, g->n_p3, g->p1_start, g->dp1,
                       g->p2_start, g->dp2, g->p3_start, g->dp3, sp.mass, sp.kB, tab);
    if ((rc = launch_status("bgk_table_kernel"))) return rc;
  }
  QLParams q;
  memset(&q, 0, sizeof(q));
  q.n1 = g->n_q1;
  q.n2 = g->n_q2;
  q.P = sp.P;
  q.n_p1 = g->n_p1;
  q.n_p2 = g->n_p2;
  q.n_p3 = g->n_p3;
  q.u = st->u;
  q.w = st->w;
  q.un = st->un;
  q.out = st->out;
  q.a = st->a;
  q.b = st->b;
  q.cdt = st->cdt;
  q.inv_dq1 = sp.inv_dq1;
  q.inv_dq2 = sp.inv_dq2;
  q.inv_tau = sp.inv_tau;
  q.source = sp.source;
  q.p_dim = sp.p_dim;
  q.p1s = g->p1_start;
  q.dp1 = g->dp1;
  q.p2s = g->p2_start;
  q.dp2 = g->dp2;
  q.p3s = g->p3_start;
  q.dp3 = g->dp3;
  q.tab = tab;
  const bool w_is_u = st->w == st->u;
  int mode = QL_GENERAL;
  if (w_is_u && st->a != 0.0) mode = QL_WU_COMB;
  else if (st->a == 0.0 && st->b == 1.0) mode = w_is_u ? QL_WU_PLAIN : QL_W_PLAIN;
  const int pd = sp.source ? sp.p_dim : 0;
  const long groups = (sp.P + 3) / 4;
  q.nseg = ql_segments(g->n_q2, groups, ql_slots<T>(g->n_q1, ph->recon_q, mode, pd));
  const int nb = (int)(groups * q.nseg);
  ql_launch<T>(q, ph->recon_q, nb, mode, pd, s);
  if ((rc = launch_status("qline_kernel"))) return rc;
  if (st->mom_out) return ql_moments<T>(g, sp, st->out, st->mom_out, part, s);
  return BOLTX_OK;
}

int check_physics(const boltx_physics* ph) {
  if (!ph) return fail(BOLTX_EINVAL, "physics is NULL");
  if (ph->recon_q < 0 || ph->recon_q > 3 || ph->recon_p < 0 || ph->recon_p > 3)
    return fail(BOLTX_EINVAL, "reconstruction out of range (q %d, p %d)", ph->recon_q, ph->recon_p);
  if (ph->riemann_q < 0 || ph->riemann_q > 1 || ph->riemann_p < 0 || ph->riemann_p > 1)
    return fail(BOLTX_EINVAL, "riemann solver out of range");
  if (ph->p_dim < 1 || ph->p_dim > 3) return fail(BOLTX_EINVAL, "p_dim = %d", ph->p_dim);
  if (ph->precision != BOLTX_FP64 && ph->precision != BOLTX_FP32_BULK)
    return fail(BOLTX_EINVAL, "precision = %d", ph->precision);
  if (ph->source_enabled && !(ph->tau > 0.0)) return fail(BOLTX_EINVAL, "tau must be > 0");
  if (ph->mass == 0.0) return fail(BOLTX_EINVAL, "mass must be non-zero");
  return BOLTX_OK;
}

}  // namespace bx

using namespace bx;

extern "C" int boltx_version(void) { return BOLTX_VERSION; }

extern "C" const char* boltx_last_error(void) { return g_err; }

// Argument checks of one stage (everything but the workspace).
static int check_stage(const boltx_grid* grid, const boltx_physics* phys, const boltx_stage* st) {
  int rc = check_grid(grid);
  if (rc) return rc;
  if ((rc = check_physics(phys))) return rc;
  if (!st || !st->u || !st->w || !st->out) return fail(BOLTX_EINVAL, "stage pointers missing");
  if (st->out == st->u) return fail(BOLTX_EINVAL, "out must not alias the stencil input u");
  if (st->a != 0.0 && !st->un) return fail(BOLTX_EINVAL, "un is NULL but a != 0");
  if (phys->source_enabled && !st->mom_in) return fail(BOLTX_EINVAL, "mom_in is NULL");
  if (phys->p_advection && (!st->fields_c_p1 || !st->fields_c_p2 || !st->fields_c_p3))
    return fail(BOLTX_EINVAL, "fields are NULL with p_advection on");
  if (st->currents_out && grid->layout != BOLTX_LAYOUT_REFERENCE)
    return fail(BOLTX_EUNSUPPORTED, "currents_out needs the reference layout");
  const int need = recon_reach(phys->recon_q) + 1;
  if ((grid->n_q1 > 1 && grid->n_q1 < need) || (grid->n_q2 > 1 && grid->n_q2 < need))
    return fail(BOLTX_EINVAL, "q-grid smaller than the stencil reach");
  if (ghost_rows(grid) > 0 && grid->n_ghost < need)
    return fail(BOLTX_EINVAL, "n_ghost = %d is below the stencil reach %d", grid->n_ghost, need);
  return BOLTX_OK;
}

template <typename T>
static int launch_currents(const StageParams& sp, int rq, double* J, hipStream_t s);

// The currents of sp.u (boltx_currents) on the stream, for the stage paths
// whose kernel does not form them itself.
static int currents_of(const StageParams& sp, const boltx_physics* phys, double* J,
                       hipStream_t s) {
  if (phys->precision == BOLTX_FP64) return launch_currents<double>(sp, phys->recon_q, J, s);
  return launch_currents<float>(sp, phys->recon_q, J, s);
}

// One stage with the given workspace (the caller's, or the scratch array).
static int fvm_stage_impl(const boltx_grid* grid, const boltx_physics* phys,
                          const boltx_stage* st, void* workspace, hipStream_t s) {
  StageParams sp = base_params(grid);
  sp.riemann_q = phys->riemann_q;
  sp.riemann_p = phys->riemann_p;
  sp.source = phys->source_enabled;
  sp.p_dim = phys->p_dim;
  sp.mass = phys->mass;
  sp.charge = phys->charge;
  sp.kB = phys->boltzmann_constant;
  sp.inv_tau = phys->source_enabled ? 1.0 / phys->tau : 0.0;
  sp.a = st->a;
  sp.b = st->b;
  sp.cdt = st->cdt;
  sp.u = st->u;
  sp.w = st->w;
  sp.un = st->un;
  sp.out = st->out;
  sp.mom_in = st->mom_in;
  sp.mom_out = st->mom_out;
  sp.fp1 = st->fields_c_p1;
  sp.fp2 = st->fields_c_p2;
  sp.fp3 = st->fields_c_p3;
  const bool padv = phys->p_advection != 0;
  int rc;
  if (grid->layout == BOLTX_LAYOUT_Q1_LINES) {
    if (padv) return fail(BOLTX_EUNSUPPORTED, "p-advection needs the reference layout");
    if (!workspace) return fail(BOLTX_EINVAL, "the q1-line layout needs the workspace");
    if (phys->precision == BOLTX_FP64) return ql_stage<double>(grid, phys, st, workspace, sp, s);
    return ql_stage<float>(grid, phys, st, workspace, sp, s);
  }
  if (workspace && qc_applies(grid)) {
    // the slab kernels accumulate the currents of u with the moments
    QCParams qp;
    sp.Jout = st->currents_out;
    qp.sp = sp;
    qp.nslab = qc_nslab(grid);
    qp.perm = (const int*)workspace;
    qp.jpack = qp.perm + (size_t)qp.nslab * 64;
    if (phys->precision == BOLTX_FP64)
      return launch_qc<double>(qp, phys->recon_q, phys->recon_p, padv, s);
    return launch_qc<float>(qp, phys->recon_q, phys->recon_p, padv, s);
  }
  if (st->currents_out && (rc = currents_of(sp, phys, st->currents_out, s))) return rc;
  if (phys->precision == BOLTX_FP64)
    return launch_stage_rq<double>(sp, phys->recon_q, phys->recon_p, padv, s);
  return launch_stage_rq<float>(sp, phys->recon_q, phys->recon_p, padv, s);
}

extern "C" int boltx_fvm_stage(const boltx_grid* grid, const boltx_physics* phys,
                               const boltx_stage* st, void* stream) {
  int rc = check_stage(grid, phys, st);
  if (rc) return rc;
  return fvm_stage_impl(grid, phys, st, st->workspace, (hipStream_t)stream);
}

static Coeff coeff_of(const boltx_coeff& c) {
  Coeff k;
  k.ptr = c.ptr;
  k.sp = (long)c.stride_p;
  k.s1 = (long)c.stride_q1;
  k.s2 = (long)c.stride_q2;
  return k;
}

extern "C" int boltx_fvm_stage_coeffs(const boltx_grid* grid, const boltx_physics* phys,
                                      const boltx_stage* st, const boltx_coeffs* co,
                                      void* stream) {
  if (!co) return fail(BOLTX_EINVAL, "coeffs is NULL");
  if (!grid) return fail(BOLTX_EINVAL, "grid is NULL");
  if (grid->layout != BOLTX_LAYOUT_REFERENCE)
    return fail(BOLTX_EUNSUPPORTED, "caller-supplied terms need the reference layout");
  const bool user_cp = co->c_p1.ptr || co->c_p2.ptr || co->c_p3.ptr;
  if (user_cp && !(co->c_p1.ptr && co->c_p2.ptr && co->c_p3.ptr))
    return fail(BOLTX_EINVAL, "c_p1, c_p2 and c_p3 are given together");
  if (user_cp && (!phys || !phys->p_advection))
    return fail(BOLTX_EINVAL, "c_p needs p_advection");
  if (co->source.ptr && co->tau.ptr)
    return fail(BOLTX_EINVAL, "tau applies to the built-in source, not to a caller's source");
  if (st && st->currents_out && (co->c_q1.ptr || co->c_q2.ptr))
    return fail(BOLTX_EINVAL, "currents_out takes its face states from the built-in C_q");
  // with the caller's source no moment block of u is read, and with the
  // caller's C_p no fields: check the stage as if both were there
  boltx_stage chk = *st;
  if (co->source.ptr && !chk.mom_in) chk.mom_in = (const double*)co->source.ptr;
  if (user_cp)
    chk.fields_c_p1 = chk.fields_c_p2 = chk.fields_c_p3 = co->c_p1.ptr;
  int rc = check_stage(grid, phys, &chk);
  if (rc) return rc;
  StageParams sp = base_params(grid);
  sp.riemann_q = phys->riemann_q;
  sp.riemann_p = phys->riemann_p;
  // sp.source turns the built-in BGK on; a caller's source replaces it
  sp.source = phys->source_enabled && !co->source.ptr;
  sp.p_dim = phys->p_dim;
  sp.mass = phys->mass;
  sp.charge = phys->charge;
  sp.kB = phys->boltzmann_constant;
  sp.inv_tau = phys->source_enabled && !co->tau.ptr ? 1.0 / phys->tau : 0.0;
  sp.a = st->a;
  sp.b = st->b;
  sp.cdt = st->cdt;
  sp.u = st->u;
  sp.w = st->w;
  sp.un = st->un;
  sp.out = st->out;
  sp.mom_in = st->mom_in;
  sp.mom_out = st->mom_out;
  sp.fp1 = st->fields_c_p1;
  sp.fp2 = st->fields_c_p2;
  sp.fp3 = st->fields_c_p3;
  sp.cq1 = coeff_of(co->c_q1);
  sp.cq2 = coeff_of(co->c_q2);
  sp.cp1 = coeff_of(co->c_p1);
  sp.cp2 = coeff_of(co->c_p2);
  sp.cp3 = coeff_of(co->c_p3);
  sp.src = coeff_of(co->source);
  sp.tau = coeff_of(co->tau);
  // the built-in BGK needs the Maxwellian of u; a caller's source does not
  if (co->source.ptr) sp.mom_in = nullptr;
  hipStream_t s = (hipStream_t)stream;
  const bool padv = phys->p_advection != 0;
  if (st->currents_out && (rc = currents_of(sp, phys, st->currents_out, s))) return rc;
  if (phys->precision == BOLTX_FP64)
    return launch_stage_rq<double>(sp, phys->recon_q, phys->recon_p, padv, s);
  return launch_stage_rq<float>(sp, phys->recon_q, phys->recon_p, padv, s);
}

extern "C" int boltx_moments(const boltx_grid* grid, const boltx_physics* phys, const void* f,
                             double* mom, void* workspace, void* stream) {
  int rc = check_grid(grid);
  if (rc) return rc;
  if ((rc = check_physics(phys))) return rc;
  if (!f || !mom) return fail(BOLTX_EINVAL, "moments pointers missing");
  StageParams sp = base_params(grid);
  sp.u = f;
  sp.mom_out = mom;
  sp.mass = phys->mass;
  sp.kB = phys->boltzmann_constant;
  sp.p_dim = phys->p_dim;
  hipStream_t s = (hipStream_t)stream;
  if (grid->layout == BOLTX_LAYOUT_Q1_LINES) {
    if (!workspace) return fail(BOLTX_EINVAL, "the q1-line layout needs the workspace");
    double* part = (double*)((char*)workspace + ql_table_bytes(grid, phys->precision));
    if (phys->precision == BOLTX_FP64) return ql_moments<double>(grid, sp, f, mom, part, s);
    return ql_moments<float>(grid, sp, f, mom, part, s);
  }
  if (phys->precision == BOLTX_FP64)
    hipLaunchKernelGGL(moments_kernel<double>, dim3(sp.Q), dim3(kBlock), 0, s, sp);
  else
    hipLaunchKernelGGL(moments_kernel<float>, dim3(sp.Q), dim3(kBlock), 0, s, sp);
  return launch_status("moments_kernel");
}

extern "C" int boltx_fill_ghosts(const boltx_grid* grid, const boltx_physics* phys, void* f,
                                 const void* const dirichlet[4], void* stream) {
  int rc = check_grid(grid);
  if (rc) return rc;
  if ((rc = check_physics(phys))) return rc;
  const long rows = ghost_rows(grid);
  if (rows == 0) return BOLTX_OK;
  if (!f) return fail(BOLTX_EINVAL, "f is NULL");
  GhostParams g;
  memset(&g, 0, sizeof(g));
  g.n_q1 = grid->n_q1;
  g.n_q2 = grid->n_q2;
  g.n_p1 = grid->n_p1;
  g.n_p2 = grid->n_p2;
  g.P = grid->n_p1 * grid->n_p2 * grid->n_p3;
  g.Q = grid->n_q1 * grid->n_q2;
  g.ng = grid->n_ghost;
  g.gh1 = grid->bc[0] != BOLTX_BC_PERIODIC;
  g.rows = (int)rows;
  g.p1s = grid->p1_start;
  g.dp1 = grid->dp1;
  g.p2s = grid->p2_start;
  g.dp2 = grid->dp2;
  for (int k = 0; k < 4; ++k) {
    g.bc[k] = grid->bc[k];
    g.dir[k] = dirichlet ? dirichlet[k] : nullptr;
    if ((g.bc[k] == BOLTX_BC_DIRICHLET || g.bc[k] == BOLTX_BC_MIRROR_DIRICHLET) && !g.dir[k])
      return fail(BOLTX_EINVAL, "dirichlet values of side %d are NULL", k);
  }
  const long n = rows * g.P;
  const unsigned nb = (unsigned)((n + 255) / 256);
  hipStream_t s = (hipStream_t)stream;
  if (phys->precision == BOLTX_FP64)
    hipLaunchKernelGGL(fill_ghosts_kernel<double>, dim3(nb), dim3(256), 0, s, g, (double*)f);
  else
    hipLaunchKernelGGL(fill_ghosts_kernel<float>, dim3(nb), dim3(256), 0, s, g, (float*)f);
  return launch_status("fill_ghosts_kernel");
}

extern "C" size_t boltx_workspace_size(const boltx_grid* grid, const boltx_physics* phys) {
  if (check_grid(grid) || check_physics(phys)) return 0;
  if (grid->layout == BOLTX_LAYOUT_Q1_LINES) return ql_workspace_bytes(grid, phys->precision);
  if (!qc_applies(grid)) return 0;
  return qc_workspace_bytes(qc_nslab(grid));
}

// The sign-grouped p order of the reference layout, written into the first
// qc_workspace_bytes() of the workspace (every entry, padding included).
static int init_slab_perm(const boltx_grid* grid, void* workspace, hipStream_t s) {
  const int nslab = qc_nslab(grid);
  int* perm = (int*)workspace;
  hipLaunchKernelGGL(slab_perm_kernel, dim3(1), dim3(64), 0, s, perm, perm + (size_t)nslab * 64,
                     grid->n_p1, grid->n_p2, grid->n_p3, grid->p1_start, grid->dp1,
                     grid->p2_start, grid->dp2);
  return launch_status("slab_perm_kernel");
}

extern "C" int boltx_workspace_init(const boltx_grid* grid, const boltx_physics* phys,
                                    void* workspace, void* stream) {
  int rc = check_grid(grid);
  if (rc) return rc;
  if ((rc = check_physics(phys))) return rc;
  if (grid->layout == BOLTX_LAYOUT_Q1_LINES || !qc_applies(grid)) return BOLTX_OK;
  if (!workspace) return fail(BOLTX_EINVAL, "workspace is NULL");
  return init_slab_perm(grid, workspace, (hipStream_t)stream);
}

// ---- the stage without a caller workspace ------------------------------------
// Device memory owned by the library for boltx_fvm_stage_noalloc: one copy per
// device (each GPU the library runs on holds it until the process ends),
// shared by every call on that device.  In host code ``scratch`` names a
// host-side stand-in: kernels get the device address from scratch_address().
static __device__ __align__(256) unsigned char scratch[BOLTX_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;

// This device's address of ``scratch``, looked up on the first call with that
// device current and kept per device.
static int scratch_address(void** out) {
  static std::mutex mu;
  static void* addr[kMaxDevices] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(BOLTX_ELAUNCH, "hipGetDevice failed");
  if (dev < 0 || dev >= kMaxDevices)
    return fail(BOLTX_EUNSUPPORTED, "device %d is beyond the %d the scratch table holds", dev,
                kMaxDevices);
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(BOLTX_ELAUNCH, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                  hipGetErrorString(e));
    addr[dev] = p;
  }
  *out = addr[dev];
  return BOLTX_OK;
}

extern "C" int boltx_fvm_stage_noalloc(const boltx_grid* grid, const boltx_physics* phys,
                                       const boltx_stage* st, void* stream) {
  int rc = check_stage(grid, phys, st);
  if (rc) return rc;
  if (st->workspace)
    return fail(BOLTX_EINVAL, "workspace must be NULL: this entry uses the library's scratch");
  const size_t need = boltx_workspace_size(grid, phys);
  if (need > (size_t)BOLTX_SCRATCH_BYTES)
    return fail(BOLTX_EUNSUPPORTED,
                "the problem needs %zu bytes of workspace, above the %zu of the scratch array",
                need, (size_t)BOLTX_SCRATCH_BYTES);
  hipStream_t s = (hipStream_t)stream;
  void* ws = nullptr;
  if (need > 0) {
    if ((rc = scratch_address(&ws))) return rc;
    // the previous call may have left another problem's p order here
    if (grid->layout == BOLTX_LAYOUT_REFERENCE && (rc = init_slab_perm(grid, ws, s))) return rc;
  }
  return fvm_stage_impl(grid, phys, st, ws, s);
}

template <typename T>
static int launch_currents(const StageParams& sp, int rq, double* J, hipStream_t s) {
  switch (rq) {
    case RC_PC: hipLaunchKernelGGL((currents_kernel<T, RC_PC>), dim3(sp.Q), dim3(kBlock), 0, s, sp, J); break;
    case RC_MINMOD: hipLaunchKernelGGL((currents_kernel<T, RC_MINMOD>), dim3(sp.Q), dim3(kBlock), 0, s, sp, J); break;
    case RC_PPM: hipLaunchKernelGGL((currents_kernel<T, RC_PPM>), dim3(sp.Q), dim3(kBlock), 0, s, sp, J); break;
    case RC_WENO5: hipLaunchKernelGGL((currents_kernel<T, RC_WENO5>), dim3(sp.Q), dim3(kBlock), 0, s, sp, J); break;
    default: return fail(BOLTX_EINVAL, "recon_q = %d", rq);
  }
  return launch_status("currents_kernel");
}

extern "C" int boltx_currents(const boltx_grid* grid, const boltx_physics* phys, const void* u,
                              double* J, void* stream) {
  int rc = check_grid(grid);
  if (rc) return rc;
  if ((rc = check_physics(phys))) return rc;
  if (!u || !J) return fail(BOLTX_EINVAL, "currents pointers missing");
  if (grid->layout != BOLTX_LAYOUT_REFERENCE)
    return fail(BOLTX_EUNSUPPORTED, "currents need the reference layout");
  StageParams sp = base_params(grid);
  sp.u = u;
  sp.charge = phys->charge;
  hipStream_t s = (hipStream_t)stream;
  if (phys->precision == BOLTX_FP64) return launch_currents<double>(sp, phys->recon_q, J, s);
  return launch_currents<float>(sp, phys->recon_q, J, s);
}

extern "C" int boltx_fdtd_step(const boltx_grid* grid, double* yee, const double* J, double dt,
                               double eps, double mu, void* stream) {
  int rc = check_grid(grid);
  if (rc) return rc;
  if (!yee || !J) return fail(BOLTX_EINVAL, "fdtd pointers missing");
  Grid2 g{grid->n_q1, grid->n_q2};
  const int Q = g.n1 * g.n2;
  hipStream_t s = (hipStream_t)stream;
  const int nb = (Q + 255) / 256;
  hipLaunchKernelGGL(fdtd_e_kernel, dim3(nb), dim3(256), 0, s, g, yee, J, grid->dq1, grid->dq2,
                     dt, eps, mu);
  if ((rc = launch_status("fdtd_e_kernel"))) return rc;
  hipLaunchKernelGGL(fdtd_b_kernel, dim3(nb), dim3(256), 0, s, g, yee, grid->dq1, grid->dq2, dt);
  return launch_status("fdtd_b_kernel");
}

extern "C" int boltx_yee_to_cell_centered(const boltx_grid* grid, const double* yee, double* cc,
                                          void* stream) {
  int rc = check_grid(grid);
  if (rc) return rc;
  if (!yee || !cc || yee == cc) return fail(BOLTX_EINVAL, "yee/cc pointers invalid");
  Grid2 g{grid->n_q1, grid->n_q2};
  const int Q = g.n1 * g.n2;
  hipLaunchKernelGGL(yee_to_cc_kernel, dim3((Q + 255) / 256), dim3(256), 0, (hipStream_t)stream, g,
                     yee, cc);
  return launch_status("yee_to_cc_kernel");
}

extern "C" int boltx_check_finite(const boltx_grid* grid, const double* mom, int32_t* flag,
                                  void* stream) {
  int rc = check_grid(grid);
  if (rc) return rc;
  if (!mom || !flag) return fail(BOLTX_EINVAL, "check_finite pointers missing");
  const int Q = grid->n_q1 * grid->n_q2;
  hipLaunchKernelGGL(check_finite_kernel, dim3((Q + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     mom, Q, flag);
  return launch_status("check_finite_kernel");
}

// ---- adjoint stage -----------------------------------------------------------
extern "C" size_t boltx_adjoint_workspace_size(const boltx_grid* grid, const boltx_physics*) {
  if (!grid) return 0;
  // moment weights [5][Q], then the q1-line reduction's chunk sums [ADJ_PCH][6][Q]
  return (size_t)(5 + 6 * ADJ_PCH) * grid->n_q1 * grid->n_q2 * sizeof(double);
}

extern "C" int boltx_adjoint_stage(const boltx_grid* grid, const boltx_physics* phys,
                                   const boltx_adjoint* adj, void* stream) {
  int rc = check_grid(grid);
  if (rc) return rc;
  if ((rc = check_physics(phys))) return rc;
  if (!adj || !adj->u || !adj->lam || !adj->out) return fail(BOLTX_EINVAL, "adjoint pointers missing");
  if (adj->out == adj->u || adj->out == adj->lam)
    return fail(BOLTX_EINVAL, "out must not alias u or lam");
  if (phys->precision != BOLTX_FP64) return fail(BOLTX_EUNSUPPORTED, "the adjoint is FP64 only");
  if (phys->p_advection) return fail(BOLTX_EUNSUPPORTED, "the adjoint needs fields off");
  if (phys->source_enabled && (!adj->mom_u || !adj->workspace))
    return fail(BOLTX_EINVAL, "the source needs mom_u and the workspace");
  const int need = recon_reach(phys->recon_q) + 1;
  if ((grid->n_q1 > 1 && grid->n_q1 < need) || (grid->n_q2 > 1 && grid->n_q2 < need))
    return fail(BOLTX_EINVAL, "q-grid smaller than the stencil reach");
  if (ghost_rows(grid) > 0)
    return fail(BOLTX_EUNSUPPORTED, "the adjoint needs periodic q boundaries");
  const AdjParams ap = adj_params(grid, phys, adj);
  // every limit is checked before the first launch: a rejected call queues
  // no device work
  if ((long)ap.n1 * ap.P >= (1L << 31) || ap.n2 > 65535)
    return fail(BOLTX_EUNSUPPORTED, "the adjoint grid takes n_q1*P < 2^31 and n_q2 <= 65535");
  hipStream_t s = (hipStream_t)stream;
  const long Q = (long)ap.n1 * ap.n2;
  if (ap.source) {
    if (ap.layout) {
      double* part = ap.gq + 5 * Q;
      const unsigned nb1 = (unsigned)((ap.n1 + 63) / 64) * ap.n2;
      hipLaunchKernelGGL(adj_bgk_reduce_lines_kernel, dim3(nb1, ADJ_PCH), dim3(64 * ADJ_RW), 0, s,
                         ap, part);
      if ((rc = launch_status("adj_bgk_reduce_lines_kernel"))) return rc;
      hipLaunchKernelGGL(adj_reduce_finish_kernel, dim3((unsigned)((Q + 255) / 256)), dim3(256), 0,
                         s, ap, (const double*)part);
    } else {
      hipLaunchKernelGGL(adj_bgk_reduce_kernel, dim3((unsigned)Q), dim3(256), 0, s, ap);
    }
    if ((rc = launch_status("adj_bgk_reduce_kernel"))) return rc;
  }
  if (ap.layout && phys->recon_q == RC_WENO5 && (ap.n1 == 64 || ap.n1 == 128)) {
    // q1-line march: one wave per (p-cell, q2 segment); segments of at least
    // 16 rows, enough of them for about 1024 workgroups
    const long groups = (ap.P + 3) / 4;
    const long nseg = std::max(1L, std::min((1024 + groups - 1) / groups, (long)(ap.n2 / 16)));
    const dim3 grid((unsigned)(groups * nseg)), block(256);
    if (ap.n1 == 64) hipLaunchKernelGGL(adj_line_kernel<1>, grid, block, 0, s, ap, (int)nseg);
    else hipLaunchKernelGGL(adj_line_kernel<2>, grid, block, 0, s, ap, (int)nseg);
    return launch_status("adj_line_kernel");
  }
  const dim3 nb((unsigned)(((long)ap.n1 * ap.P + 255) / 256), (unsigned)ap.n2);
  switch (phys->recon_q) {
    case RC_PC: hipLaunchKernelGGL(adj_stage_kernel<RC_PC>, nb, dim3(256), 0, s, ap); break;
    case RC_MINMOD:
      hipLaunchKernelGGL(adj_stage_kernel<RC_MINMOD>, nb, dim3(256), 0, s, ap);
      break;
    case RC_PPM: hipLaunchKernelGGL(adj_stage_kernel<RC_PPM>, nb, dim3(256), 0, s, ap); break;
    default: hipLaunchKernelGGL(adj_stage_kernel<RC_WENO5>, nb, dim3(256), 0, s, ap);
  }
  return launch_status("adj_stage_kernel");
}
