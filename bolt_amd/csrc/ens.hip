// C-ABI of the ensemble step (declarations in include/boltx.h): chunk-list
// planning on the host, the chunk stage launch (stage and partial moments of
// its output), the batched moment-block / Maxwellian-table launch, and the
// moment launch for a given state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cstdint>
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
  std::vector<EnsChunk> chunks;  // in launch order: largest first
  std::vector<int2> fin_blocks;  // finalize workgroups: (instance, first q1 cell)
  long pooled = 0, tab = 0, mom = 0, part = 0;
  size_t lds = 0;                // dynamic LDS bytes of the chunk kernel
  // byte offsets inside the workspace
  size_t o_inst = 0, o_chunks = 0, o_fin = 0, o_mom = 0, o_part = 0, o_tab = 0, o_flag = 0,
         total = 0;
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
  if (t->chunk_cells < 1 || t->batch_cells < 1)
    return fail(BOLTX_EINVAL, "ensemble tuning: chunk_cells and batch_cells must be positive");
  return BOLTX_OK;
}

// Chunk list and workspace map of an ensemble under a tuning.
Plan make_plan(int n, const boltx_ens_member* m, const boltx_ens_tuning* t) {
  Plan pl;
  const int esz = m[0].phys.precision == BOLTX_FP64 ? 8 : 4;
  const int vw = 16 / esz;  // cells per 16-byte vector
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
    I.rn_p1 = 1.0f / float(g.n_p1);
    I.rn_p2 = 1.0f / float(g.n_p2);
    I.P = g.n_p1 * g.n_p2 * g.n_p3;
    I.p_dim = ph.p_dim;
    I.source = ph.source_enabled;
    I.off_f = pl.pooled;
    pl.pooled += (long)I.P * I.n1;
    // table blocks start on 16-byte boundaries
    I.off_tab = (pl.tab + vw - 1) / vw * vw;
    pl.tab = I.off_tab + (long)(g.n_p1 + g.n_p2 + g.n_p3) * I.n1;
    // items of two 16-byte vectors where the rows allow them
    I.w = I.n1 % (2 * vw) == 0 && I.off_f % vw == 0 ? 2 * vw : 1;
    I.off_mom = pl.mom;
    pl.mom += (long)MOM_ROWS * I.n1;
    // whole q1 lines per chunk and per batch; a batch is at most
    // ENS_THREADS items of I.w cells
    const int lc = std::min(I.P, std::max(1, t->chunk_cells / I.n1));
    I.batch = std::max(1, std::min({t->batch_cells / I.n1, ENS_THREADS / (I.n1 / I.w), lc}));
    I.nchunk = (I.P + lc - 1) / lc;
    I.off_part = pl.part;
    pl.part += (long)I.nchunk * 5 * I.n1;
    for (int c = 0; c < I.nchunk; ++c)
      pl.chunks.push_back({i, c, c * lc, std::min(I.P, (c + 1) * lc)});
    for (int q0 = 0; q0 < I.n1; q0 += ENS_FIN_CELLS) pl.fin_blocks.push_back(make_int2(i, q0));
    const size_t lds = esz == 8 ? ens_chunk_lds<double>(I.n1, I.batch)
                                : ens_chunk_lds<float>(I.n1, I.batch);
    pl.lds = std::max(pl.lds, lds);
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
  // launch the largest chunks first, so the small ones fill the tail
  std::stable_sort(pl.chunks.begin(), pl.chunks.end(), [&](const EnsChunk& a, const EnsChunk& b) {
    return (long)(a.p1 - a.p0) * pl.inst[a.inst].n1 > (long)(b.p1 - b.p0) * pl.inst[b.inst].n1;
  });
  // workspace map
  size_t o = 0;
  pl.o_inst = o;
  o = align256(o + pl.inst.size() * sizeof(EnsInst));
  pl.o_chunks = o;
  o = align256(o + pl.chunks.size() * sizeof(EnsChunk));
  pl.o_fin = o;
  o = align256(o + pl.fin_blocks.size() * sizeof(int2));
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

// The chunk kernel reads and writes 16-byte vectors of the pooled arrays and
// of the workspace tables.
int check_aligned(const void* p, const char* what) {
  if ((uintptr_t)p % 16)
    return fail(BOLTX_EINVAL, "ensemble: %s must be 16-byte aligned", what);
  return BOLTX_OK;
}

template <typename T>
int finalize(const Plan& pl, char* ws, hipStream_t s) {
  hipLaunchKernelGGL(ens_finalize_kernel<T>, dim3((unsigned)pl.fin_blocks.size()),
                     dim3(ENS_FIN_CELLS * ENS_FIN_SUBS), 0, s, (const EnsInst*)(ws + pl.o_inst),
                     (const int2*)(ws + pl.o_fin), (const double*)(ws + pl.o_part),
                     (double*)(ws + pl.o_mom), (T*)(ws + pl.o_tab), (int*)(ws + pl.o_flag));
  return launch_status("ens_finalize_kernel");
}

template <typename T>
int moments_and_tables(const Plan& pl, char* ws, const void* f, hipStream_t s) {
  hipLaunchKernelGGL(ens_moments_kernel<T>, dim3((unsigned)pl.chunks.size()), dim3(256), 0, s,
                     (const EnsInst*)(ws + pl.o_inst), (const EnsChunk*)(ws + pl.o_chunks),
                     (const T*)f, (double*)(ws + pl.o_part));
  int rc = launch_status("ens_moments_kernel");
  if (rc) return rc;
  return finalize<T>(pl, ws, s);
}

template <typename T>
int stage(const Plan& pl, char* ws, int recon, const void* u, const void* w, const void* un,
          void* out, double a, double b, double cscale, hipStream_t s) {
  EnsStageArgs ea;
  ea.inst = (const EnsInst*)(ws + pl.o_inst);
  ea.chunks = (const EnsChunk*)(ws + pl.o_chunks);
  ea.u = u;
  ea.w = w;
  ea.un = un;
  ea.out = out;
  ea.tab = ws + pl.o_tab;
  ea.part = (double*)(ws + pl.o_part);
  ea.a = a;
  ea.b = b;
  ea.cscale = cscale;
  ea.use_un = a != 0.0;
  ea.b_one = b == 1.0;
  const dim3 grid((unsigned)pl.chunks.size()), block(ENS_THREADS);
  const unsigned lds = (unsigned)pl.lds;
  const int form = w == u && a == 0.0 && b == 1.0 ? ENS_FORM_FIRST
                   : w == u && a != 0.0           ? ENS_FORM_COMBINE
                                                  : ENS_FORM_GENERAL;
  auto go = [&](auto rc_, auto form_) {
    hipLaunchKernelGGL((ens_chunk_kernel<T, decltype(rc_)::value, decltype(form_)::value>), grid,
                       block, lds, s, ea);
  };
  auto by_form = [&](auto rc_) {
    using F0 = std::integral_constant<int, ENS_FORM_FIRST>;
    using F1 = std::integral_constant<int, ENS_FORM_COMBINE>;
    using F2 = std::integral_constant<int, ENS_FORM_GENERAL>;
    if (form == ENS_FORM_FIRST) go(rc_, F0());
    else if (form == ENS_FORM_COMBINE) go(rc_, F1());
    else go(rc_, F2());
  };
  switch (recon) {
    case RC_PC: by_form(std::integral_constant<int, RC_PC>()); break;
    case RC_MINMOD: by_form(std::integral_constant<int, RC_MINMOD>()); break;
    case RC_PPM: by_form(std::integral_constant<int, RC_PPM>()); break;
    default: by_form(std::integral_constant<int, RC_WENO5>());
  }
  int rc = launch_status("ens_chunk_kernel");
  if (rc) return rc;
  return finalize<T>(pl, ws, s);
}

template <typename T>
int steps(const Plan& pl, char* ws, int recon, int stepper, void* f, void* u1, void* u2,
          int n_steps, hipStream_t s) {
  int rc = BOLTX_OK;
  for (int k = 0; k < n_steps && !rc; ++k) {
    if (stepper == BOLTX_RK2) {
      // fvm_operator.py:25-33: u1 = f + dt/2 L(f); f = f + dt L(u1)
      if ((rc = stage<T>(pl, ws, recon, f, f, nullptr, u1, 0.0, 1.0, 0.5, s))) break;
      rc = stage<T>(pl, ws, recon, u1, f, nullptr, f, 0.0, 1.0, 1.0, s);
    } else {
      if ((rc = stage<T>(pl, ws, recon, f, f, nullptr, u1, 0.0, 1.0, 1.0, s))) break;
      if ((rc = stage<T>(pl, ws, recon, u1, u1, f, u2, 0.75, 0.25, 1.0, s))) break;
      rc = stage<T>(pl, ws, recon, u2, u2, f, f, 1.0 / 3.0, 2.0 / 3.0, 1.0, s);
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
  return make_plan(n, members, tuning).total;
}

extern "C" int boltx_ensemble_map(int n, const boltx_ens_member* members,
                                  const boltx_ens_tuning* tuning, size_t* moments_offset,
                                  size_t* flag_offset) {
  int rc = check_members(n, members);
  if (rc || (rc = check_tuning(tuning))) return rc;
  const Plan pl = make_plan(n, members, tuning);
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
  if ((rc = check_aligned(f, "f")) || (rc = check_aligned(workspace, "the workspace"))) return rc;
  const Plan pl = make_plan(n, members, tuning);
  char* ws = (char*)workspace;
  hipStream_t s = (hipStream_t)stream;
  struct Piece {
    size_t off;
    const void* src;
    size_t bytes;
  } pieces[] = {{pl.o_inst, pl.inst.data(), pl.inst.size() * sizeof(EnsInst)},
                {pl.o_chunks, pl.chunks.data(), pl.chunks.size() * sizeof(EnsChunk)},
                {pl.o_fin, pl.fin_blocks.data(), pl.fin_blocks.size() * sizeof(int2)}};
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
  if ((rc = check_aligned(f, "f")) || (rc = check_aligned(u1, "u1")) ||
      (u2 && (rc = check_aligned(u2, "u2"))) || (rc = check_aligned(workspace, "the workspace")))
    return rc;
  const Plan pl = make_plan(n, members, tuning);
  char* ws = (char*)workspace;
  hipStream_t s = (hipStream_t)stream;
  const int recon = members[0].phys.recon_q;
  if (members[0].phys.precision == BOLTX_FP64)
    return steps<double>(pl, ws, recon, stepper, f, u1, u2, n_steps, s);
  return steps<float>(pl, ws, recon, stepper, f, u1, u2, n_steps, s);
}
