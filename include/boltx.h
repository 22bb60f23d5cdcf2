/* C-ABI of the MI355X (gfx950) nonlinear FVM time step.
 *
 * The reference (QuazarTech/Bolt) is pure Python over ArrayFire; it has no FFI.
 * Each entry point below replaces the whole-array ArrayFire work of one
 * reference function (cited per function), and is called only by the Python
 * mirror of the reference interface (bolt_amd.nonlinear_solver), never by user
 * code.  Conventions:
 *   - plain C: sizes, spacings and device pointers; no allocation of
 *     persistent memory inside the library;
 *   - every call is asynchronous on the given HIP stream (a hipStream_t passed
 *     as void*, NULL = the null stream) and re-entrant;
 *   - return 0 on success, a negative boltx_status otherwise; the message is
 *     in boltx_last_error() (thread-local).
 *
 * Layout of a distribution function: interior q-cells (periodic q is handled
 * by modular indexing, so a periodic axis has no ghost zones), P = n_p1*n_p2*n_p3
 * with p-index i1 + n_p1*(i2 + n_p2*i3) (p1 fastest), and either
 *   BOLTX_LAYOUT_REFERENCE: C-order [n_q2][n_q1][P], the byte order of the
 *     reference's (P, S, Q1, Q2) column-major af.Array for S = 1 without its
 *     ghost rows; or
 *   BOLTX_LAYOUT_Q1_LINES: C-order [n_q2][P][n_q1] (q1 fastest, n_q1 <= 256),
 *     used for fields-off problems: every q1 line of one p-cell is contiguous.
 * FP64 storage is double, FP32-bulk storage is float.
 * Non-periodic q boundaries (grid->bc, reference layout only): every f-sized
 * array holds G = 2*n_ghost*(n_q2 if q1 is non-periodic) + 2*n_ghost*(n_q1 if
 * q2 is non-periodic) ghost q-rows of P values after its Q interior rows:
 * left [c][i2] (q1 = c - n_ghost), right [c][i2] (q1 = n_q1 + c), bottom
 * [c][i1] (q2 = c - n_ghost), top [c][i1] (q2 = n_q2 + c), c = 0..n_ghost-1,
 * the blocks of periodic axes left out.  boltx_fill_ghosts fills them; the
 * stage kernels read them as the stencil neighbours beyond the boundary.
 * A moment block is an FP64 array [BOLTX_MOM_ROWS][Q] (Q = n_q1*n_q2): rows
 * 0-4 are density, mom_v1_bulk, mom_v2_bulk, mom_v3_bulk and energy, each
 * already multiplied by dp3*dp2*dp1; rows 5-11 are the BGK parameters derived
 * from them (n (m/2 pi k T)^(p_dim/2), u1, u2, u3, T, n, m/(2kT)).  Every entry
 * point that writes a moment block writes all of it.
 * Cell-centred fields are FP64 arrays [6][Q]: E1, E2, E3, B1, B2, B3.
 */
#ifndef BOLTX_H
#define BOLTX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BOLTX_VERSION 1
#define BOLTX_MOM_ROWS 12

enum boltx_recon {
  BOLTX_RECON_PIECEWISE_CONSTANT = 0,
  BOLTX_RECON_MINMOD = 1,
  BOLTX_RECON_PPM = 2,
  BOLTX_RECON_WENO5 = 3
};
enum boltx_riemann { BOLTX_RIEMANN_UPWIND = 0, BOLTX_RIEMANN_LLF = 1 };
enum boltx_precision { BOLTX_FP64 = 0, BOLTX_FP32_BULK = 1 };
enum boltx_layout { BOLTX_LAYOUT_REFERENCE = 0, BOLTX_LAYOUT_Q1_LINES = 1 };
/* q boundary of one side (boundaries.py:415-577; 'none' is a singleton axis
 * marked periodic: it has no neighbours). */
enum boltx_bc {
  BOLTX_BC_PERIODIC = 0,
  BOLTX_BC_MIRROR = 1,
  BOLTX_BC_DIRICHLET = 2,
  BOLTX_BC_MIRROR_DIRICHLET = 3
};
enum boltx_status {
  BOLTX_OK = 0,
  BOLTX_EINVAL = -1,
  BOLTX_ELAUNCH = -2,
  BOLTX_EUNSUPPORTED = -3
};

typedef struct boltx_grid {
  int32_t n_q1, n_q2;       /* interior q-cells */
  int32_t n_p1, n_p2, n_p3; /* p-cells per axis */
  int32_t n_species;        /* 1 */
  int32_t layout;           /* enum boltx_layout of every f-sized array */
  double dq1, dq2;
  double dp1, dp2, dp3;
  double p1_start, p2_start, p3_start;
  int32_t n_ghost;          /* ghost width of the non-periodic q axes */
  int32_t bc[4];            /* enum boltx_bc of the left, right, bottom, top sides;
                               a periodic axis is periodic on both sides */
} boltx_grid;

typedef struct boltx_physics {
  int32_t recon_q, recon_p;     /* enum boltx_recon */
  int32_t riemann_q, riemann_p; /* enum boltx_riemann */
  int32_t source_enabled;       /* BGK collision source on */
  int32_t p_advection;          /* FVM advection in p (fields enabled) */
  int32_t p_dim;                /* 1, 2 or 3 */
  int32_t precision;            /* enum boltx_precision */
  double mass, charge, boltzmann_constant;
  double tau;                   /* constant collision time */
} boltx_physics;

typedef struct boltx_stage {
  const void *u;   /* stage input, read through the stencils */
  const void *w;   /* base of the update, read pointwise (may equal u or un) */
  const void *un;  /* step-start state, read pointwise when a != 0 */
  void *out;       /* stage output; must not alias u (may alias w or un) */
  double a, b, cdt; /* out = a*un + b*(w + cdt*L(u)) */
  const double *mom_in;  /* moment block of u (needed when the source is on) */
  double *mom_out;       /* moment block of out, or NULL */
  const double *fields_c_p1; /* [6][Q] fields seen by C_p1 (p_advection) */
  const double *fields_c_p2; /* [6][Q] fields seen by C_p2 */
  const double *fields_c_p3; /* [6][Q] fields seen by C_p3 */
  void *workspace;  /* boltx_workspace_size() bytes set up by boltx_workspace_init(),
                       or NULL; stages sharing a workspace must be stream-ordered */
  double *currents_out; /* [3][Q] FP64 or NULL: the currents of u as boltx_currents
                           computes them (reference layout only), written by the
                           stage itself where its kernel forms the face states */
} boltx_stage;

/* One Runge-Kutta stage of the FVM step: out = a*un + b*(w + cdt*L(u)) where
 * L(u) = -div_q(C_q u) + BGK(u) - div_p(C_p u) with the reference's
 * reconstruction and upwind / LLF fluxes, fused with the moments of out.
 * Replaces df_dt_fvm (bolt/lib/nonlinear/finite_volume/df_dt_fvm.py:33-339)
 * together with one stage of timestep_fvm (fvm_operator.py:22-33) and the
 * BGK source (bolt/src/nonrelativistic_boltzmann/collision_operator.py:70-128).
 * In the q1-line layout the q2 extent is marched in segments of about 16
 * rows; the environment variable BOLTX_QL_SEGMENTS=<count>, read at each
 * call, fixes the segment count instead (a tuning and test knob: results are
 * bit-identical for every count). */
int boltx_fvm_stage(const boltx_grid *grid, const boltx_physics *phys,
                    const boltx_stage *stage, void *stream);

/* boltx_fvm_stage with no workspace argument and no allocation.  It has no
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  stage->workspace must be NULL; the workspace lives
 * in a static device array of BOLTX_SCRATCH_BYTES that the library holds on
 * every GPU it runs on until the process ends.  A problem whose
 * boltx_workspace_size() exceeds it is rejected with BOLTX_EUNSUPPORTED.
 * Every call on one GPU shares that GPU's array: calls whose device work
 * overlaps (other streams, other host threads) race, so order them on one
 * stream.  Outputs equal those of boltx_fvm_stage bit for bit. */
#define BOLTX_SCRATCH_BYTES (64u << 20)
int boltx_fvm_stage_noalloc(const boltx_grid *grid, const boltx_physics *phys,
                            const boltx_stage *stage, void *stream);

/* Caller-supplied terms of the stage: user callbacks (C_q, C_p, source,
 * tau of the reference's physical_system plugins) evaluated on the device.
 * A coefficient is an FP64 array read at p-cell p of interior q-cell
 * (i1, i2) as ptr[p*stride_p + i1*stride_q1 + i2*stride_q2] (element
 * strides; 0 broadcasts, so a callback's broadcast result is passed
 * without materialising it); ptr = NULL selects the built-in term. */
typedef struct boltx_coeff {
  const double *ptr;
  int64_t stride_p, stride_q1, stride_q2;
} boltx_coeff;

typedef struct boltx_coeffs {
  /* C_q1 at the left q1-face of cell (i1, i2), read for i1 = 0..n_q1 (the
   * C_q callback at q1_left_center, q2_left_center: df_dt_fvm.py:67-71) */
  boltx_coeff c_q1;
  /* C_q2 at the bottom q2-face, read for i2 = 0..n_q2 (at q1_center_bot,
   * q2_center_bot: df_dt_fvm.py:73-77) */
  boltx_coeff c_q2;
  /* C_p1, C_p2, C_p3 at the left, bottom and back p-faces of every p-cell
   * (the C_p callback at p1_left / p2_bottom / p3_back: df_dt_fvm.py:251-267);
   * all three or none, with p_advection on */
  boltx_coeff c_p1, c_p2, c_p3;
  /* an additive source term evaluated on the stage input, applied in place
   * of the built-in BGK (the source callback: df_dt_fvm.py:91-97) */
  boltx_coeff source;
  /* the built-in BGK's collision time per cell (params.tau(q1, q2, p1, p2,
   * p3): collision_operator.py:119-128), instead of phys->tau */
  boltx_coeff tau;
} boltx_coeffs;

/* boltx_fvm_stage with caller-supplied terms (reference layout; the generic
 * one-workgroup-per-q-cell kernel).  Replaces df_dt_fvm
 * (df_dt_fvm.py:33-339) for any advection_terms / source module, inside one
 * stage of timestep_fvm (fvm_operator.py:22-33).  With the caller's source,
 * mom_in is not read; with the caller's C_p, the fields are not read.  The
 * caller keeps every coefficient array alive and unchanged until the stage
 * has run on the stream; the indices above must stay inside them. */
int boltx_fvm_stage_coeffs(const boltx_grid *grid, const boltx_physics *phys,
                           const boltx_stage *stage, const boltx_coeffs *coeffs,
                           void *stream);

/* The moment block of f.  Replaces compute_moments
 * (bolt/lib/nonlinear/compute_moments.py:7-59) for density, mom_v1_bulk,
 * mom_v2_bulk, mom_v3_bulk and energy (moments.py:5-22), plus the BGK
 * parameters (collision_operator.py:102-116). */
int boltx_moments(const boltx_grid *grid, const boltx_physics *phys, const void *f,
                  double *mom, void *workspace, void *stream);

/* Fill the ghost q-rows of f (layout above) for the non-periodic sides, as
 * communicate_f + apply_bcs_f leave them before a RHS evaluation
 * (communicate.py:8-38 with the 'ghosted' DMDA axes of
 * nonlinear_solver.py:177-201, then boundaries.py:415-577): start from zero
 * (the DMDA never writes a ghosted boundary zone); mirror sides copy the
 * mirror image of the first n_ghost interior cells with the normal velocity
 * reversed (p1 for left/right, p2 for bottom/top; apply_mirror_bcs_f_cartesian,
 * boundaries.py:227-311); dirichlet sides then take dirichlet[side] where the
 * characteristic enters (C_q1 = p1 > 0 on the left, < 0 on the right, C_q2 = p2
 * likewise; apply_dirichlet_bcs_f, boundaries.py:147-225).  dirichlet[side]
 * holds that side's ghost block ([c][i_other][P], storage type) of the user's
 * f_<side> values, and may be NULL for non-dirichlet sides. */
int boltx_fill_ghosts(const boltx_grid *grid, const boltx_physics *phys, void *f,
                      const void *const dirichlet[4], void *stream);

/* Bytes of device workspace boltx_fvm_stage and boltx_moments use for this
 * problem (required in the q1-line layout, optional in the reference layout,
 * where it holds the sign-grouped p order), and its one-time set-up.  The
 * reference has no counterpart: ArrayFire allocates its temporaries itself;
 * here the caller owns all device memory. */
size_t boltx_workspace_size(const boltx_grid *grid, const boltx_physics *phys);
int boltx_workspace_init(const boltx_grid *grid, const boltx_physics *phys,
                         void *workspace, void *stream);

/* Currents J1, J2, J3 ([3][Q] FP64) from the upwind face states of u:
 * J1 = charge * mom_v1_bulk(f_left), J2 = charge * mom_v2_bulk(f_bot),
 * J3 = charge * mom_v3_bulk(f).  Replaces df_dt_fvm.py:214-224. */
int boltx_currents(const boltx_grid *grid, const boltx_physics *phys,
                   const void *u, double *J, void *stream);

/* One FDTD leapfrog on the periodic Yee grid, in place on yee ([6][Q] FP64:
 * E1(i,j+1/2), E2(i+1/2,j), E3(i+1/2,j+1/2), B1(i+1/2,j), B2(i,j+1/2),
 * B3(i,j)): E += dt/(mu eps) curl B - dt J / eps, then B -= dt curl E.
 * Replaces fdtd (bolt/lib/nonlinear/fields/electrodynamic/fdtd_explicit.py:141-177). */
int boltx_fdtd_step(const boltx_grid *grid, double *yee, const double *J,
                    double dt, double eps, double mu, void *stream);

/* Yee grid -> cell centres ([6][Q] each).  Replaces
 * yee_grid_to_cell_centered_grid (bolt/lib/nonlinear/fields/fields.py:527-557). */
int boltx_yee_to_cell_centered(const boltx_grid *grid, const double *yee,
                               double *cc, void *stream);

/* Sets *flag (device int) to 1 if any moment density is not finite; a
 * non-finite f makes its density sum non-finite.  Replaces check_divergence
 * (bolt/lib/nonlinear/timestep.py:28-38) without a host sync. */
int boltx_check_finite(const boltx_grid *grid, const double *mom, int32_t *flag,
                       void *stream);

/* Adjoint of one stage (FP64, fields off, periodic q: a grid with ghost
 * q-rows is rejected with BOLTX_EUNSUPPORTED).  For the stage
 * out = a*un + b*(w + cdt*L(u)) and an adjoint lam of its output, the caller
 * passes c = b*cdt and gets out = a*base + a2*base2 + b*lam + c*J(u)^T lam,
 * J = dL/du, and, if tau_grad is non-NULL, tau_grad[q] = c * sum_p lam *
 * dL/dtau for each q-cell (the BGK source's dependence on its collision
 * time).  base and base2 may be NULL and may alias out; out must not alias u
 * or lam.  The reference has
 * no adjoint: this differentiates df_dt_fvm.py:33-97 with the BGK source
 * collision_operator.py:70-128 and the moments moments.py:5-22. */
typedef struct boltx_adjoint {
  const void *u;          /* forward stage input */
  const double *mom_u;    /* moment block of u (needed when the source is on) */
  const void *lam;        /* adjoint of the stage output */
  const void *base;       /* or NULL */
  void *out;
  double a, b, c;
  double *tau_grad;       /* [Q] or NULL */
  void *workspace;        /* boltx_adjoint_workspace_size() bytes */
  const void *base2;      /* or NULL: a second term a2*base2 (SSP-RK3 reverses its */
  double a2;              /* first stage into lam/3 + 3/4 lam_2 + ...) */
} boltx_adjoint;

int boltx_adjoint_stage(const boltx_grid *grid, const boltx_physics *phys,
                        const boltx_adjoint *adj, void *stream);
size_t boltx_adjoint_workspace_size(const boltx_grid *grid, const boltx_physics *phys);

/* Ensembles: independent members that are one-dimensional in q (periodic
 * q1, n_q2 = 1, fields off, n_q1 <= 256), each with its own grids, physics
 * and time step, advanced together.  Every member's f lives in the q1-line
 * layout [P][n_q1] inside pooled device arrays (member i at the element
 * offset boltx_ensemble_layout() reports); all members share the
 * reconstruction and the precision.  Each stage of the whole ensemble is one
 * launch of a workgroup per (member, chunk of chunk_cells cells in whole
 * q1 lines) that forms the stage and the FP64 partial moments of its output,
 * then one batched moment-block / Maxwellian-table launch.  The pooled
 * arrays and the workspace must be 16-byte aligned.  The reference runs such
 * ensembles one instance at a time: this replaces that loop over
 * timestep_fvm (fvm_operator.py:6-50) for many small problems. */
enum boltx_stepper { BOLTX_RK2 = 0, BOLTX_SSP_RK3 = 1 };

typedef struct boltx_ens_member {
  boltx_grid grid;
  boltx_physics phys;
  double dt;
} boltx_ens_member;

typedef struct boltx_ens_tuning {
  int32_t chunk_cells; /* target cells per chunk (rounded down to whole q1 lines, at least one);
                          the moment sums are grouped by chunk, so results depend on it */
  int32_t batch_cells; /* target cells a chunk workgroup stages in LDS at once (whole lines);
                          results do not depend on it */
} boltx_ens_tuning;

/* Element offsets of the members in the pooled f-sized arrays (offsets[n],
 * may be NULL); returns the pooled element count, or -1. */
long long boltx_ensemble_layout(int n, const boltx_ens_member *members, long long *offsets);
/* Bytes of device workspace (work lists, moment blocks, partial moments,
 * Maxwellian tables, divergence flag) for this ensemble and tuning; 0 on
 * invalid input. */
size_t boltx_ensemble_workspace_size(int n, const boltx_ens_member *members,
                                     const boltx_ens_tuning *tuning);
/* Byte offsets inside the workspace of the moment blocks (member i's
 * [BOLTX_MOM_ROWS][n_q1] block after those of members 0..i-1) and of the
 * int32 divergence flag. */
int boltx_ensemble_map(int n, const boltx_ens_member *members, const boltx_ens_tuning *tuning,
                       size_t *moments_offset, size_t *flag_offset);
/* Writes the work lists into the workspace (synchronising the stream once)
 * and computes the moment blocks and tables of the pooled state f. */
int boltx_ensemble_init(int n, const boltx_ens_member *members, const boltx_ens_tuning *tuning,
                        const void *f, void *workspace, void *stream);
/* n_steps steps of every member (stepper: enum boltx_stepper), f in place,
 * u1 and u2 pooled scratch states (u2 unused by RK2). */
int boltx_ensemble_step(int n, const boltx_ens_member *members, const boltx_ens_tuning *tuning,
                        int stepper, void *f, void *u1, void *u2, void *workspace, int n_steps,
                        void *stream);

int boltx_version(void);
const char *boltx_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* BOLTX_H */
