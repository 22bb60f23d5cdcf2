// C-ABI of the ensemble step (declarations in include/boltx.h): work-list
// planning on the host, the persistent stage launch, and the batched moment
// and Maxwellian-table launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/boltx.h"
#include "bx_status.h"
#include "ensemble.h"

namespace bx {
namespace {

struct Plan {
  std::vector<EnsInst> inst;
  std::vector<EnsTile> tiles;
  std::vector<int> wg_begin;
  std::vector<EnsChunk> chunks;
  long pooled = 0, tab = 0, mom = 0, part = 0;
  int groups = 0;
  // byte offsets inside the workspace
  size_t o_inst = 0, o_tiles = 0, o_wg = 0, o_chunks = 0, o_mom = 0, o_part = 0, o_tab = 0,
         o_flag = 0, total = 0;
};

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

int check_members(int n, const boltx_ens_member* m) {
  if (n <= 0 || !m) return fail(BOLTX_EINVAL, "ensemble: no members");
  for (int i = 0; i < n; ++i) {
    const boltx_grid& g = m[i].grid;
    const boltx_physics& ph = m[i].phys;
    int rc = check_grid(&g);
    if (!rc) rc = check_physics(&ph);
    if (rc) {
      char msg[400];
      snprintf(msg, sizeof(msg), "%s", boltx_last_error());
      return fail(rc, "ensemble member %d: %s", i, msg);
    }
    if (g.n_q2 != 1 || g.n_q1 > 256 || g.n_q1 < 5)
      return fail(BOLTX_EUNSUPPORTED, "ensemble member %d: needs N_q2 = 1 and 5 <= N_q1 <= 256", i);
    if (ph.p_advection) return fail(BOLTX_EUNSUPPORTED, "ensemble member %d: fields must be off", i);
    if (ph.recon_q != m[0].phys.recon_q || ph.precision != m[0].phys.precision)
      return fail(BOLTX_EINVAL, "ensemble members must share reconstruction and precision");
    if (!(m[i].dt > 0.0)) return fail(BOLTX_EINVAL, "ensemble member %d: dt <= 0", i);
  }
  return BOLTX_OK;
}

int check_tuning(const boltx_ens_tuning* t) {
  if (!t) return fail(BOLTX_EINVAL, "ensemble: tuning is NULL");
  if (t->waves_per_group != 1 && t->waves_per_group != 2 && t->waves_per_group != 4)
    return fail(BOLTX_EINVAL, "waves_per_group must be 1, 2 or 4");
  if (t->cells_per_task < 1 || t->groups_per_cu < 1 || t->groups_per_cu > 64 || t->chunk_cells < 1)
    return fail(BOLTX_EINVAL, "ensemble tuning out of range");
  return BOLTX_OK;
}

int cu_count() {
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
    return 256;
  return n;
}

// Work lists and workspace map of an ensemble under a tuning.
Plan make_plan(int n, const boltx_ens_member* m, const boltx_ens_tuning* t, int ncu) {
  Plan pl;
  const int esz = m[0].phys.precision == BOLTX_FP64 ? 8 : 4;
  // members in order of decreasing line length keep one line shape per
  // stretch of the work list
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return m[a].grid.n_q1 > m[b].grid.n_q1; });
  pl.inst.resize(n);
  for (int i = 0; i < n; ++i) {
    const boltx_grid& g = m[i].grid;
    const boltx_physics& ph = m[i].phys;
    EnsInst& I = pl.inst[i];
    memset(&I, 0, sizeof(I));
    I.n1 = g.n_q1;
    I.n_p1 = g.n_p1;
    I.n_p2 = g.n_p2;
    I.n_p3 = g.n_p3;
    I.P = g.n_p1 * g.n_p2 * g.n_p3;
    I.p_dim = ph.p_dim;
    I.source = ph.source_enabled;
    I.lanes = I.n1 <= 64 ? 64 / I.n1 : 1;
    I.off_f = pl.pooled;
    pl.pooled += (long)I.P * I.n1;
    I.off_tab = pl.tab;
    pl.tab += (long)(g.n_p1 + g.n_p2 + g.n_p3) * I.n1;
    I.off_mom = pl.mom;
    pl.mom += (long)MOM_ROWS * I.n1;
    const int pc = std::max(1, t->chunk_cells / I.n1);
    I.nchunk = (I.P + pc - 1) / pc;
    I.off_part = pl.part;
    pl.part += (long)I.nchunk * 5 * I.n1;
    for (int c = 0; c < I.nchunk; ++c)
      pl.chunks.push_back({i, c, c * pc, std::min(I.P, (c + 1) * pc)});
    I.inv_dq = 1.0 / g.dq1;
    I.cdt = m[i].dt;
    I.inv_tau = ph.source_enabled ? 1.0 / ph.tau : 0.0;
    I.p1s = g.p1_start;
    I.dp1 = g.dp1;
    I.p2s = g.p2_start;
    I.dp2 = g.dp2;
    I.p3s = g.p3_start;
    I.dp3 = g.dp3;
    I.measure = g.dp1 * g.dp2 * g.dp3;
    I.mass = ph.mass;
    I.kB = ph.boltzmann_constant;
  }
  // wave tasks of about cells_per_task cells (whole waves of packed lines)
  std::vector<long> cost;
  for (int oi = 0; oi < n; ++oi) {
    const int i = order[oi];
    const EnsInst& I = pl.inst[i];
    int np = std::max(1, t->cells_per_task / I.n1);
    if (I.n1 <= 64) np = (np + I.lanes - 1) / I.lanes * I.lanes;
    for (int p0 = 0; p0 < I.P; p0 += np) {
      const int k = std::min(np, I.P - p0);
      pl.tiles.push_back({i, p0, k, 0});
      // a wave pays for whole 64-lane rows
      const long rows = I.n1 <= 64 ? (k + I.lanes - 1) / I.lanes : (long)k * ((I.n1 + 63) / 64);
      cost.push_back(rows);
    }
  }
  // contiguous ranges of equal cost for the persistent workgroups
  long total = 0;
  for (long c : cost) total += c;
  const int G = std::max(1, std::min<int>((int)pl.tiles.size(), t->groups_per_cu * ncu));
  pl.groups = G;
  pl.wg_begin.assign(G + 1, (int)pl.tiles.size());
  pl.wg_begin[0] = 0;
  long acc = 0;
  int g = 1;
  for (size_t k = 0; k < pl.tiles.size() && g < G; ++k) {
    acc += cost[k];
    while (g < G && acc * G >= total * g) pl.wg_begin[g++] = (int)k + 1;
  }
  for (; g < G; ++g) pl.wg_begin[g] = (int)pl.tiles.size();
  // workspace map
  size_t o = 0;
  pl.o_inst = o;
  o = align256(o + pl.inst.size() * sizeof(EnsInst));
  pl.o_tiles = o;
  o = align256(o + pl.tiles.size() * sizeof(EnsTile));
  pl.o_wg = o;
  o = align256(o + pl.wg_begin.size() * sizeof(int));
  pl.o_chunks = o;
  o = align256(o + pl.chunks.size() * sizeof(EnsChunk));
  pl.o_mom = o;
  o = align256(o + (size_t)pl.mom * sizeof(double));
  pl.o_part = o;
  o = align256(o + (size_t)pl.part * sizeof(double));
  pl.o_tab = o;
  o = align256(o + (size_t)pl.tab * esz);
  pl.o_flag = o;
  o = align256(o + sizeof(int));
  pl.total = o;
  return pl;
}

template <typename T>
int moments_and_tables(const Plan& pl, char* ws, const void* f, hipStream_t s) {
  const EnsInst* inst = (const EnsInst*)(ws + pl.o_inst);
  hipLaunchKernelGGL(ens_moments_kernel<T>, dim3((unsigned)pl.chunks.size()), dim3(256), 0, s,
                     inst, (const EnsChunk*)(ws + pl.o_chunks), (const T*)f,
                     (double*)(ws + pl.o_part));
  int rc = launch_status("ens_moments_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(ens_finalize_kernel<T>, dim3((unsigned)pl.inst.size(), ENS_TGROUPS), dim3(256), 0, s,
                     inst, (const double*)(ws + pl.o_part), (double*)(ws + pl.o_mom),
                     (T*)(ws + pl.o_tab), (int*)(ws + pl.o_flag));
  return launch_status("ens_finalize_kernel");
}

template <typename T>
int stage(const Plan& pl, char* ws, int recon, int wpg, const void* u, const void* w,
          const void* un, void* out, double a, double b, double cscale, hipStream_t s) {
  EnsStageArgs ea;
  ea.inst = (const EnsInst*)(ws + pl.o_inst);
  ea.tiles = (const EnsTile*)(ws + pl.o_tiles);
  ea.wg_begin = (const int*)(ws + pl.o_wg);
  ea.u = u;
  ea.w = w;
  ea.un = un;
  ea.out = out;
  ea.tab = ws + pl.o_tab;
  ea.a = a;
  ea.b = b;
  ea.cscale = cscale;
  ea.use_un = a != 0.0;
  ea.b_one = b == 1.0;
  const dim3 grid(pl.groups), block(64 * wpg);
  switch (recon) {
    case RC_PC: hipLaunchKernelGGL((ens_stage_kernel<T, RC_PC>), grid, block, 0, s, ea); break;
    case RC_MINMOD:
      hipLaunchKernelGGL((ens_stage_kernel<T, RC_MINMOD>), grid, block, 0, s, ea);
      break;
    case RC_PPM: hipLaunchKernelGGL((ens_stage_kernel<T, RC_PPM>), grid, block, 0, s, ea); break;
    default: hipLaunchKernelGGL((ens_stage_kernel<T, RC_WENO5>), grid, block, 0, s, ea);
  }
  int rc = launch_status("ens_stage_kernel");
  if (rc) return rc;
  return moments_and_tables<T>(pl, ws, out, s);
}

template <typename T>
int steps(const Plan& pl, char* ws, int recon, int wpg, int stepper, void* f, void* u1, void* u2,
          int n_steps, hipStream_t s) {
  int rc = BOLTX_OK;
  for (int k = 0; k < n_steps && !rc; ++k) {
    if (stepper == BOLTX_RK2) {
      // fvm_operator.py:25-33: u1 = f + dt/2 L(f); f = f + dt L(u1)
      if ((rc = stage<T>(pl, ws, recon, wpg, f, f, nullptr, u1, 0.0, 1.0, 0.5, s))) break;
      rc = stage<T>(pl, ws, recon, wpg, u1, f, nullptr, f, 0.0, 1.0, 1.0, s);
    } else {
      if ((rc = stage<T>(pl, ws, recon, wpg, f, f, nullptr, u1, 0.0, 1.0, 1.0, s))) break;
      if ((rc = stage<T>(pl, ws, recon, wpg, u1, u1, f, u2, 0.75, 0.25, 1.0, s))) break;
      rc = stage<T>(pl, ws, recon, wpg, u2, u2, f, f, 1.0 / 3.0, 2.0 / 3.0, 1.0, s);
    }
  }
  return rc;
}

}  // namespace
}  // namespace bx

using namespace bx;

extern "C" long long boltx_ensemble_layout(int n, const boltx_ens_member* members,
                                           long long* offsets) {
  if (check_members(n, members)) return -1;
  long long o = 0;
  for (int i = 0; i < n; ++i) {
    if (offsets) offsets[i] = o;
    o += (long long)members[i].grid.n_q1 * members[i].grid.n_p1 * members[i].grid.n_p2 *
         members[i].grid.n_p3;
  }
  return o;
}

extern "C" size_t boltx_ensemble_workspace_size(int n, const boltx_ens_member* members,
                                                const boltx_ens_tuning* tuning) {
  if (check_members(n, members) || check_tuning(tuning)) return 0;
  return make_plan(n, members, tuning, cu_count()).total;
}

extern "C" int boltx_ensemble_map(int n, const boltx_ens_member* members,
                                  const boltx_ens_tuning* tuning, size_t* moments_offset,
                                  size_t* flag_offset) {
  int rc = check_members(n, members);
  if (rc || (rc = check_tuning(tuning))) return rc;
  const Plan pl = make_plan(n, members, tuning, cu_count());
  if (moments_offset) *moments_offset = pl.o_mom;
  if (flag_offset) *flag_offset = pl.o_flag;
  return BOLTX_OK;
}

extern "C" int boltx_ensemble_init(int n, const boltx_ens_member* members,
                                   const boltx_ens_tuning* tuning, const void* f,
                                   void* workspace, void* stream) {
  int rc = check_members(n, members);
  if (rc || (rc = check_tuning(tuning))) return rc;
  if (!f || !workspace) return fail(BOLTX_EINVAL, "ensemble: f or workspace is NULL");
  const Plan pl = make_plan(n, members, tuning, cu_count());
  char* ws = (char*)workspace;
  hipStream_t s = (hipStream_t)stream;
  struct Piece {
    size_t off;
    const void* src;
    size_t bytes;
  } pieces[] = {{pl.o_inst, pl.inst.data(), pl.inst.size() * sizeof(EnsInst)},
                {pl.o_tiles, pl.tiles.data(), pl.tiles.size() * sizeof(EnsTile)},
                {pl.o_wg, pl.wg_begin.data(), pl.wg_begin.size() * sizeof(int)},
                {pl.o_chunks, pl.chunks.data(), pl.chunks.size() * sizeof(EnsChunk)}};
  for (const Piece& pc : pieces)
    if (hipMemcpyAsync(ws + pc.off, pc.src, pc.bytes, hipMemcpyHostToDevice, s) != hipSuccess)
      return fail(BOLTX_ELAUNCH, "ensemble: copy of the work lists failed");
  const int zero = 0;
  if (hipMemcpyAsync(ws + pl.o_flag, &zero, sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess)
    return fail(BOLTX_ELAUNCH, "ensemble: flag reset failed");
  // the host vectors must outlive the copies
  if (hipStreamSynchronize(s) != hipSuccess) return fail(BOLTX_ELAUNCH, "ensemble: sync failed");
  if (members[0].phys.precision == BOLTX_FP64) return moments_and_tables<double>(pl, ws, f, s);
  return moments_and_tables<float>(pl, ws, f, s);
}

extern "C" int boltx_ensemble_step(int n, const boltx_ens_member* members,
                                   const boltx_ens_tuning* tuning, int stepper, void* f, void* u1,
                                   void* u2, void* workspace, int n_steps, void* stream) {
  int rc = check_members(n, members);
  if (rc || (rc = check_tuning(tuning))) return rc;
  if (!f || !u1 || !workspace || (stepper != BOLTX_RK2 && !u2))
    return fail(BOLTX_EINVAL, "ensemble: state or workspace pointers missing");
  if (stepper != BOLTX_RK2 && stepper != BOLTX_SSP_RK3)
    return fail(BOLTX_EINVAL, "ensemble: stepper %d", stepper);
  const Plan pl = make_plan(n, members, tuning, cu_count());
  char* ws = (char*)workspace;
  hipStream_t s = (hipStream_t)stream;
  const int recon = members[0].phys.recon_q, wpg = tuning->waves_per_group;
  if (members[0].phys.precision == BOLTX_FP64)
    return steps<double>(pl, ws, recon, wpg, stepper, f, u1, u2, n_steps, s);
  return steps<float>(pl, ws, recon, wpg, stepper, f, u1, u2, n_steps, s);
}
