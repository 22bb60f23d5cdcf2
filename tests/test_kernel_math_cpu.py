"""The stage kernel's per-cell code (bolt_amd/csrc/fvm_stage.h), run on the
CPU through tests/hostcheck, against the FP64 oracle: one RK stage from the
same state.  Tolerances: FP64 <= 1e-12 relative L-inf of f (the per-step
parity bar); FP32-bulk <= 1e-5 relative L2 against the FP64 oracle."""
import numpy as np
import pytest
import torch

from oracle import bolt_oracle as bo

from tests.cases import case_2d2v, case_c1, case_2d3v, oracle_solver
from tests.hostcheck import harness as H

RECON = ["piecewise-constant", "minmod", "ppm", "weno5"]
RIEMANN = ["upwind-flux", "lax-friedrichs"]


def _one_stage(o, dt):
    o.dt = dt
    o._communicate_f()
    un = o.f
    L = o.df_dt_fvm(un)
    return un, un + L * dt


def _rel(a, b):
    d = np.abs(a.astype(np.float64) - b)
    return d.max() / np.abs(b).max(), np.sqrt((d**2).sum() / (b**2).sum())


@pytest.mark.parametrize("rq", RECON)
@pytest.mark.parametrize("rs", RIEMANN)
def test_stage_2d2v_fp64(rq, rs):
    c = case_2d2v(seed=1, N_q=12, N_p=10, recon_q=rq, riemann=rs)
    o = oracle_solver(c)
    un, ref = _one_stage(o, c.dt)
    u = H.to_internal(o, un)
    out, mom = H.stage(o, u, u, None, 0.0, 1.0, c.dt)
    linf, _ = _rel(out, H.to_internal(o, ref))
    assert linf <= 1e-12
    # the fused moments of the output equal the oracle's moments of the output
    m_ref = H.moments_np(o, H.to_internal(o, ref))
    assert np.abs(mom - m_ref).max() <= 1e-13 * np.abs(m_ref).max()


def test_stage_weno5_smooth_full_q_resolution():
    """At 128^2 the data is so smooth that the expanded WENO indicator is
    rounding-dominated; the literal evaluation keeps parity."""
    c = case_2d2v(seed=2, N_q=128, N_p=6)
    o = oracle_solver(c)
    un, ref = _one_stage(o, c.dt)
    u = H.to_internal(o, un)
    out, _ = H.stage(o, u, u, None, 0.0, 1.0, c.dt)
    linf, _ = _rel(out, H.to_internal(o, ref))
    assert linf <= 1e-12


@pytest.mark.parametrize("rq", RECON)
def test_stage_fp32_bulk(rq):
    c = case_2d2v(seed=3, N_q=16, N_p=12, recon_q=rq)
    o = oracle_solver(c)
    un, ref = _one_stage(o, c.dt)
    u = H.to_internal(o, un)
    out, _ = H.stage(o, u, u, None, 0.0, 1.0, c.dt, precision=1)
    _, l2 = _rel(out, H.to_internal(o, ref))
    assert l2 <= 1e-6


def test_rk_combination_matches_oracle():
    """out = a*un + b*(w + cdt*L(u)) for the SSP-RK3 second stage."""
    c = case_2d2v(seed=4, N_q=10, N_p=8)
    o = oracle_solver(c)
    o.dt = c.dt
    o._communicate_f()
    un = o.f
    u1 = un + o.df_dt_fvm(un) * c.dt
    o.f = u1
    o._communicate_f()
    u1 = o.f
    ref = 0.75 * un + 0.25 * (u1 + o.df_dt_fvm(u1) * c.dt)
    a1, an = H.to_internal(o, u1), H.to_internal(o, un)
    out, _ = H.stage(o, a1, a1, an, 0.75, 0.25, c.dt)
    assert _rel(out, H.to_internal(o, ref))[0] <= 1e-12


def test_stage_1d1v_singleton_q2():
    c = case_c1(seed=0)
    o = oracle_solver(c)
    o.dt = c.dt
    o._communicate_f()
    un = o.f
    ref = un + o.df_dt_fvm(un) * (c.dt / 2)
    u = H.to_internal(o, un)
    out, _ = H.stage(o, u, u, None, 0.0, 1.0, c.dt / 2)
    assert _rel(out, H.to_internal(o, ref))[0] <= 1e-14


@pytest.mark.parametrize("rp", RECON)
def test_stage_fields_p_advection(rp):
    """Fields on, first RK stage (at_n True: no FDTD); C_p1, C_p2, C_p3 read
    the aliased cell-centred fields."""
    c = case_2d3v(seed=5, N_q=6, N_p=8, recon_p=rp, B0=0.5, stepper="RK2")
    o = oracle_solver(c)
    # give the fields some structure so every C_p term matters
    fs = o.fields_solver
    g = torch_rand_like(fs.cell_centered_EM_fields, seed=9)
    fs.cell_centered_EM_fields[:] = g  # mutate the shared (aliased) object
    o.dt = c.dt
    o._communicate_f()
    un = o.f
    ref = un + o.df_dt_fvm(un) * c.dt
    u = H.to_internal(o, un)
    cc = fs.cell_centered_EM_fields[:, 0, o.N_g1:o.N_g1 + o.N_q1, o.N_g2:o.N_g2 + o.N_q2]
    F = np.ascontiguousarray(cc.permute(0, 2, 1).reshape(6, -1).numpy())
    out, _ = H.stage(o, u, u, None, 0.0, 1.0, c.dt, fields=[F, F, F])
    assert _rel(out, H.to_internal(o, ref))[0] <= 1e-12


def torch_rand_like(t, seed):
    import torch
    g = torch.Generator().manual_seed(seed)
    return torch.rand(t.shape, generator=g, dtype=t.dtype) * 0.4 - 0.2


@pytest.mark.parametrize("rq", RECON)
@pytest.mark.parametrize("rs", RIEMANN)
@pytest.mark.parametrize("precision", [0, 1])
def test_slab_kernel_flux_matches_reference_form(rq, rs, precision):
    """The slab kernel's upwind flux difference (four shared WENO indicators,
    FP64 faces through one division, LLF through the upwind state) equals the
    reference-form evaluation (per-face indicators, reciprocal weights, literal
    LLF) to rounding, on smooth and on rough windows of both velocity signs."""
    import ctypes
    rng = np.random.default_rng(7)
    n = 4000
    x = np.arange(7) - 3.0
    smooth = 1.0 + 0.3 * np.sin(0.4 * x[None, :] + rng.uniform(0, 6, (n // 2, 1)))
    rough = rng.uniform(0.0, 1.0, (n // 2, 7))
    a7 = np.ascontiguousarray(np.vstack([smooth, rough]))
    C = rng.uniform(-5, 5, n)
    fast, ref = np.empty(n), np.empty(n)
    lib = H.lib()
    lib.host_flux_pair.restype = ctypes.c_int
    P = lambda v: v.ctypes.data_as(ctypes.c_void_p)
    lib.host_flux_pair(H.L.RECON[rq], precision, H.L.RIEMANN[rs], P(a7), P(C), P(fast),
                       P(ref), n)
    scale = np.abs(C) * np.abs(a7).max(axis=1)
    tol = 1e-14 if precision == 0 else 2e-6
    assert np.all(np.abs(fast - ref) <= tol * scale)


@pytest.mark.parametrize("rq", ["minmod", "ppm", "weno5"])
@pytest.mark.parametrize("precision", [0, 1])
def test_qline_faces_match_reference_faces(rq, precision):
    """The q1-line kernels' face arithmetic (indicator quantity per cell, the
    d-weights folded into the substencils, one division per face) against the
    oracle's faces (weno5.py:57-98, ppm.py:57-98, minmod.py:64-71) on random
    windows: 1e-13 relative in FP64, 1e-5 in FP32."""
    import ctypes
    from tests.hostcheck import harness as H
    lib = H.lib()
    lib.host_ql_faces.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_int]
    rng = np.random.default_rng(11)
    n = 4000
    y = np.ascontiguousarray(1.0 + 0.3 * rng.standard_normal((n, 5)))
    T5 = [torch.from_numpy(y[:, k].copy()) for k in range(5)]
    if rq == "weno5":
        L, R = bo.weno5_faces(*T5)
    elif rq == "ppm":
        L, R = bo.ppm_faces(*T5)
    else:
        x = torch.from_numpy(y[:, 2].copy())
        s = bo.minmod(2 * (x - T5[1]), 0.5 * (T5[3] - T5[1]), 2 * (T5[3] - x))
        L, R = x - 0.5 * s, x + 0.5 * s
    for right, ref in ((1, R.numpy()), (0, L.numpy())):
        out = np.empty(n)
        lib.host_ql_faces(H.L.RECON[rq], precision, right, y.ctypes.data, out.ctypes.data, n)
        err = np.abs(out - ref).max() / np.abs(ref).max()
        assert err <= (1e-13 if precision == 0 else 1e-5), (rq, right, err)


@pytest.mark.parametrize("precision", [0, 1])
def test_ppm_split_form_is_bitwise_ppm_faces(precision):
    """The FP32 p-sweep kernel forms its two q-faces per cell from four shared
    limited slopes (ppm_dq, then ppm_faces_dq); both faces must equal
    ppm_faces bit for bit.  The windows mix smooth data with values on a
    coarse grid, so zero slopes, flat runs, extrema and each of the three
    corrections (ppm.py:89-98) occur."""
    import ctypes
    from tests.hostcheck import harness as H
    lib = H.lib()
    lib.host_ppm_split_mismatch.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    lib.host_ppm_split_mismatch.restype = ctypes.c_int
    rng = np.random.default_rng(17)
    n = 20000
    smooth = 1.0 + 0.3 * rng.standard_normal((n, 5))
    coarse = np.round(rng.standard_normal((n, 5)) * 2) / 2
    y = np.ascontiguousarray(np.concatenate([smooth, coarse]))
    # the grid data must exercise the corrections: count windows where the
    # oracle's corrected faces differ from the unlimited parabola's
    T5 = [torch.from_numpy(y[:, k].copy()) for k in range(5)]
    L, R = bo.ppm_faces(*T5)
    plain_l = 0.5 * (T5[2] + T5[1])
    assert int((L != plain_l).sum()) > n // 10
    assert lib.host_ppm_split_mismatch(precision, y.ctypes.data, len(y)) == 0
