// CPU harness for the tests: runs the stage kernel's per-cell code
// (bolt_amd/csrc/fvm_stage.h) on host memory, one q-cell after another, so
// the kernel arithmetic and indexing can be compared with the oracle without
// a GPU.  Built host-only by tests/hostcheck/build_host.py; never shipped.
#include <cstring>
#include <vector>

#include "../../include/boltx.h"
#include "../../bolt_amd/csrc/fvm_stage.h"

using namespace bx;

static StageParams params_of(const boltx_grid* g, const boltx_physics* ph) {
  StageParams sp;
  memset(&sp, 0, sizeof(sp));
  sp.n_q1 = g->n_q1; sp.n_q2 = g->n_q2;
  sp.n_p1 = g->n_p1; sp.n_p2 = g->n_p2; sp.n_p3 = g->n_p3;
  sp.P = g->n_p1 * g->n_p2 * g->n_p3;
  sp.Q = g->n_q1 * g->n_q2;
  sp.p1s = g->p1_start; sp.p2s = g->p2_start; sp.p3s = g->p3_start;
  sp.dp1 = g->dp1; sp.dp2 = g->dp2; sp.dp3 = g->dp3;
  sp.inv_dq1 = 1.0 / g->dq1; sp.inv_dq2 = 1.0 / g->dq2;
  sp.inv_dp1 = 1.0 / g->dp1; sp.inv_dp2 = 1.0 / g->dp2; sp.inv_dp3 = 1.0 / g->dp3;
  sp.measure = g->dp3 * g->dp2 * g->dp1;
  if (ph) {
    sp.riemann_q = ph->riemann_q; sp.riemann_p = ph->riemann_p;
    sp.source = ph->source_enabled; sp.p_dim = ph->p_dim;
    sp.mass = ph->mass; sp.charge = ph->charge; sp.kB = ph->boltzmann_constant;
    sp.inv_tau = ph->source_enabled ? 1.0 / ph->tau : 0.0;
  }
  return sp;
}

template <typename T, int RQ, int RP, bool PADV>
static void run(StageParams sp) {
  std::vector<T> tab(sp.n_p1 + sp.n_p2 + sp.n_p3);
  for (int q = 0; q < sp.Q; ++q) {
    QCell<T> c;
    qcell_neighbours(sp, q, c);
    c.pref = 0.0;
    c.e1 = c.e2 = c.e3 = tab.data();
    if (sp.source) {
      const BgkCell b = qcell_bgk(sp, q);
      c.pref = maxwell_prefactor(b, sp.mass, sp.kB, sp.p_dim);
      for (int i = 0; i < sp.n_p1; ++i)
        tab[i] = T(maxwell_factor(centre(sp.p1s, i, sp.dp1), b.u1, b.T, sp.mass, sp.kB));
      for (int i = 0; i < sp.n_p2; ++i)
        tab[sp.n_p1 + i] = T(maxwell_factor(centre(sp.p2s, i, sp.dp2), b.u2, b.T, sp.mass, sp.kB));
      for (int i = 0; i < sp.n_p3; ++i)
        tab[sp.n_p1 + sp.n_p2 + i] =
            T(maxwell_factor(centre(sp.p3s, i, sp.dp3), b.u3, b.T, sp.mass, sp.kB));
      c.e1 = tab.data();
      c.e2 = tab.data() + sp.n_p1;
      c.e3 = tab.data() + sp.n_p1 + sp.n_p2;
    }
    const T* blk = (const T*)sp.u + (long)q * sp.P;
    if (PADV && !sp.cp1.ptr) qcell_fields(sp, c);
    double acc[5] = {0, 0, 0, 0, 0};
    for (int p = 0; p < sp.P; ++p) {
      double v[3];
      const T x = stage_cell<T, RQ, RP, PADV>(sp, c, p, blk, v);
      ((T*)sp.out)[(long)q * sp.P + p] = x;
      acc[0] += x; acc[1] += x * v[0]; acc[2] += x * v[1]; acc[3] += x * v[2];
      acc[4] += (0.5 * x) * ((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    }
    if (sp.mom_out)
      for (int k = 0; k < 5; ++k) sp.mom_out[k * sp.Q + q] = acc[k] * sp.measure;
  }
}

template <typename T, int RQ>
static void run_rp(StageParams sp, int rp, bool padv) {
  if (!padv) return run<T, RQ, RC_PC, false>(sp);
  switch (rp) {
    case 0: return run<T, RQ, RC_PC, true>(sp);
    case 1: return run<T, RQ, RC_MINMOD, true>(sp);
    case 2: return run<T, RQ, RC_PPM, true>(sp);
    default: return run<T, RQ, RC_WENO5, true>(sp);
  }
}

template <typename T>
static void run_rq(StageParams sp, int rq, int rp, bool padv) {
  switch (rq) {
    case 0: return run_rp<T, RC_PC>(sp, rp, padv);
    case 1: return run_rp<T, RC_MINMOD>(sp, rp, padv);
    case 2: return run_rp<T, RC_PPM>(sp, rp, padv);
    default: return run_rp<T, RC_WENO5>(sp, rp, padv);
  }
}

extern "C" int host_fvm_stage(const boltx_grid* g, const boltx_physics* ph, const boltx_stage* st) {
  StageParams sp = params_of(g, ph);
  sp.a = st->a; sp.b = st->b; sp.cdt = st->cdt;
  sp.u = st->u; sp.w = st->w; sp.un = st->un; sp.out = st->out;
  sp.mom_in = st->mom_in; sp.mom_out = st->mom_out;
  sp.fp1 = st->fields_c_p1; sp.fp2 = st->fields_c_p2; sp.fp3 = st->fields_c_p3;
  if (ph->precision == 0) run_rq<double>(sp, ph->recon_q, ph->recon_p, ph->p_advection != 0);
  else run_rq<float>(sp, ph->recon_q, ph->recon_p, ph->p_advection != 0);
  return 0;
}

static Coeff coeff_of(const boltx_coeff& c) {
  Coeff k;
  k.ptr = c.ptr; k.sp = (long)c.stride_p; k.s1 = (long)c.stride_q1; k.s2 = (long)c.stride_q2;
  return k;
}

// The generic stage with caller-supplied terms (as boltx_fvm_stage_coeffs
// sets up its parameters).
extern "C" int host_fvm_stage_coeffs(const boltx_grid* g, const boltx_physics* ph,
                                     const boltx_stage* st, const boltx_coeffs* co) {
  StageParams sp = params_of(g, ph);
  sp.a = st->a; sp.b = st->b; sp.cdt = st->cdt;
  sp.u = st->u; sp.w = st->w; sp.un = st->un; sp.out = st->out;
  sp.mom_in = st->mom_in; sp.mom_out = st->mom_out;
  sp.fp1 = st->fields_c_p1; sp.fp2 = st->fields_c_p2; sp.fp3 = st->fields_c_p3;
  sp.cq1 = coeff_of(co->c_q1); sp.cq2 = coeff_of(co->c_q2);
  sp.cp1 = coeff_of(co->c_p1); sp.cp2 = coeff_of(co->c_p2); sp.cp3 = coeff_of(co->c_p3);
  sp.src = coeff_of(co->source); sp.tau = coeff_of(co->tau);
  sp.source = ph->source_enabled && !co->source.ptr;
  sp.inv_tau = ph->source_enabled && !co->tau.ptr ? 1.0 / ph->tau : 0.0;
  if (ph->precision == 0) run_rq<double>(sp, ph->recon_q, ph->recon_p, ph->p_advection != 0);
  else run_rq<float>(sp, ph->recon_q, ph->recon_p, ph->p_advection != 0);
  return 0;
}

#include "../../bolt_amd/csrc/qcell.h"

// Flux differences of n seven-point windows two ways: the slab kernel's
// upwind evaluation (shared indicators, one division per FP64 face) and the
// reference-form flux_diff.
template <typename T, int R>
static void flux_pair(const double* a7, const double* C, double* fast, double* ref, int n, int rs) {
  for (int i = 0; i < n; ++i) {
    T a[7];
    for (int k = 0; k < 7; ++k) a[k] = T(a7[7 * i + k]);
    const T c = T(C[i]);
    fast[i] = double(upwind_flux_diff<R, T>(a, c, c > T(0)));
    ref[i] = double(flux_diff<R, T>(a, c, rs, (T*)nullptr));
  }
}

extern "C" int host_flux_pair(int recon, int precision, int riemann, const double* a7,
                              const double* C, double* fast, double* ref, int n) {
  if (precision == 0) {
    switch (recon) {
      case 0: flux_pair<double, RC_PC>(a7, C, fast, ref, n, riemann); break;
      case 1: flux_pair<double, RC_MINMOD>(a7, C, fast, ref, n, riemann); break;
      case 2: flux_pair<double, RC_PPM>(a7, C, fast, ref, n, riemann); break;
      default: flux_pair<double, RC_WENO5>(a7, C, fast, ref, n, riemann);
    }
  } else {
    switch (recon) {
      case 0: flux_pair<float, RC_PC>(a7, C, fast, ref, n, riemann); break;
      case 1: flux_pair<float, RC_MINMOD>(a7, C, fast, ref, n, riemann); break;
      case 2: flux_pair<float, RC_PPM>(a7, C, fast, ref, n, riemann); break;
      default: flux_pair<float, RC_WENO5>(a7, C, fast, ref, n, riemann);
    }
  }
  return 0;
}

#include "../../bolt_amd/csrc/adjoint.h"

// The adjoint stage's per-cell code on host memory (both passes).
extern "C" int host_adjoint_stage(const boltx_grid* g, const boltx_physics* ph,
                                  const boltx_adjoint* adj) {
  const AdjParams ap = adj_params(g, ph, adj);
  const int Q = ap.n1 * ap.n2;
  // the reduction over p split as the device kernels split it: in the
  // q1-line layout ADJ_PCH chunks of ADJ_RW interleaved streams (chunks added
  // in order), in the reference layout 256 interleaved streams added in order
  if (ap.source)
    for (int q = 0; q < Q; ++q) {
      double r[6] = {0, 0, 0, 0, 0, 0};
      if (ap.layout) {
        for (int c = 0; c < ADJ_PCH; ++c) {
          double x[6];
          adj_reduce_chunk(ap, q, c, x);
          for (int k = 0; k < 6; ++k) r[k] += x[k];
        }
      } else {
        for (int w = 0; w < 256; ++w) {
          double x[6];
          adj_reduce_sums(ap, q, w, 256, x);
          for (int k = 0; k < 6; ++k) r[k] += x[k];
        }
      }
      adj_reduce_finish(ap, q, r);
    }
  const long n = (long)Q * ap.P;
  for (long t = 0; t < n; ++t) {
    switch (ph->recon_q) {
      case 0: adj_stage_cell<RC_PC>(ap, t); break;
      case 1: adj_stage_cell<RC_MINMOD>(ap, t); break;
      case 2: adj_stage_cell<RC_PPM>(ap, t); break;
      default: adj_stage_cell<RC_WENO5>(ap, t);
    }
  }
  return 0;
}

#include "../../bolt_amd/csrc/qline.h"

// Faces of the q1-line kernels (indicator quantities per cell, one division
// per face) on n five-cell windows y[5*i .. 5*i+4]: the right (or left) face
// of the centre cell.
template <typename T, int R>
static void ql_faces(const double* y, double* out, int n, bool right) {
  for (int i = 0; i < n; ++i) {
    T a[5];
    for (int k = 0; k < 5; ++k) a[k] = T(y[5 * i + k]);
    const T ka = ql_kappa<R, T>(a[0], a[1], a[2]);
    const T kb = ql_kappa<R, T>(a[1], a[2], a[3]);
    const T kc = ql_kappa<R, T>(a[2], a[3], a[4]);
    out[i] = double(right ? ql_face<R, true, T>(a[0], a[1], a[2], a[3], a[4], ka, kb, kc)
                          : ql_face<R, false, T>(a[0], a[1], a[2], a[3], a[4], ka, kb, kc));
  }
}

extern "C" int host_ql_faces(int recon, int precision, int right, const double* y, double* out,
                             int n) {
  const bool r = right != 0;
  if (precision == 0) {
    switch (recon) {
      case 0: ql_faces<double, RC_PC>(y, out, n, r); break;
      case 1: ql_faces<double, RC_MINMOD>(y, out, n, r); break;
      case 2: ql_faces<double, RC_PPM>(y, out, n, r); break;
      default: ql_faces<double, RC_WENO5>(y, out, n, r);
    }
  } else {
    switch (recon) {
      case 0: ql_faces<float, RC_PC>(y, out, n, r); break;
      case 1: ql_faces<float, RC_MINMOD>(y, out, n, r); break;
      case 2: ql_faces<float, RC_PPM>(y, out, n, r); break;
      default: ql_faces<float, RC_WENO5>(y, out, n, r);
    }
  }
  return 0;
}

// ppm_faces against the split form the FP32 p-sweep kernel uses for its
// q-faces (ppm_dq of each cell, then ppm_faces_dq) on n five-cell windows
// y[5*i .. 5*i+4]: the number of windows whose faces differ in any bit.
template <typename T>
static int ppm_split_mismatch(const double* y, int n) {
  int bad = 0;
  for (int i = 0; i < n; ++i) {
    T a[5];
    for (int k = 0; k < 5; ++k) a[k] = T(y[5 * i + k]);
    T l0, r0, l1, r1;
    ppm_faces(a[0], a[1], a[2], a[3], a[4], l0, r0);
    ppm_faces_dq(a[1], a[2], a[3], ppm_dq(a[0], a[1], a[2]), ppm_dq(a[1], a[2], a[3]),
                 ppm_dq(a[2], a[3], a[4]), l1, r1);
    bad += (std::memcmp(&l0, &l1, sizeof(T)) != 0 || std::memcmp(&r0, &r1, sizeof(T)) != 0);
  }
  return bad;
}

extern "C" int host_ppm_split_mismatch(int precision, const double* y, int n) {
  return precision == 0 ? ppm_split_mismatch<double>(y, n) : ppm_split_mismatch<float>(y, n);
}
