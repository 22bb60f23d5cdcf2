"""Pins of the FP64 oracle (CPU): ArrayFire semantics, closed forms, the
reference's own known-answer tests and invariants, each chosen so a plausible
mistake (a flipped shift, a dropped term, a transposed axis) fails."""
import math

import numpy as np
import pytest
import torch

from oracle import bolt_oracle as bo
from tests.cases import case_2d2v, case_c1, maxwellian_case, oracle_solver

F64 = torch.float64
rng = np.random.default_rng(1234)


def T(x):
    return torch.as_tensor(np.asarray(x, dtype=np.float64))


# -- ArrayFire primitives -----------------------------------------------------
def test_af_shift_direction_and_wrap():
    x = torch.arange(10, dtype=F64).reshape(1, 1, 10, 1)
    y = bo.af_shift(x, 0, 0, 1)
    # af.shift: out[i] = in[(i - k) mod n]
    assert y.reshape(-1).tolist() == [9, 0, 1, 2, 3, 4, 5, 6, 7, 8]
    y = bo.af_shift(x, 0, 0, -2)
    assert y.reshape(-1).tolist() == [2, 3, 4, 5, 6, 7, 8, 9, 0, 1]


def test_moddims_is_column_major():
    a = rng.normal(size=(3, 4, 5, 2))
    got = bo.moddims(T(a), 6, 2, 10, 1).numpy()
    want = np.reshape(np.reshape(a, -1, order="F"), (6, 2, 10, 1), order="F")
    assert np.array_equal(got, want)


def test_p_grid_layout_p1_fastest():
    """calculate_p.py:33-57: flat index = i1 + N1 (i2 + N2 i3)."""
    p1, p2, p3 = bo.calculate_p_center([-1.0], [-2.0], [-3.0], 3, 4, 5, [0.5], [1.0], [1.5])
    idx = 2 + 3 * (1 + 4 * 3)
    assert p1.reshape(-1)[idx] == -1.0 + 2.5 * 0.5
    assert p2.reshape(-1)[idx] == -2.0 + 1.5 * 1.0
    assert p3.reshape(-1)[idx] == -3.0 + 3.5 * 1.5
    # same as numpy.meshgrid(p2, p1, p3) flattened column-major
    a1 = -1.0 + (0.5 + np.arange(3)) * 0.5
    a2 = -2.0 + (0.5 + np.arange(4)) * 1.0
    a3 = -3.0 + (0.5 + np.arange(5)) * 1.5
    g = np.meshgrid(a2, a1, a3)
    assert np.array_equal(p1.reshape(-1).numpy(), g[1].ravel(order="F"))
    assert np.array_equal(p2.reshape(-1).numpy(), g[0].ravel(order="F"))


# -- reconstruction -----------------------------------------------------------
def test_minmod_closed_form_brute_force():
    x = rng.normal(size=2000)
    y = rng.normal(size=2000)
    z = rng.normal(size=2000)
    x[:50] = 0.0
    got = bo.minmod(T(x), T(y), T(z)).numpy()
    same = (np.sign(x + 0.0) >= 0) == (np.sign(y) >= 0)
    same &= (np.sign(y) >= 0) == (np.sign(z) >= 0)
    m = np.minimum(np.minimum(abs(x), abs(y)), abs(z))
    want = np.where(same, np.where(y >= 0, m, -m), 0.0)
    assert np.array_equal(np.abs(got), np.abs(want))
    assert np.allclose(got, want, atol=0)


def _line(vals):
    return T(np.asarray(vals, dtype=np.float64)).reshape(1, 1, -1, 1)


def test_minmod_linear_exact_and_extremum_flat():
    x = _line(3.0 + 0.25 * np.arange(12))
    left, right = bo.reconstruct_minmod(x, 2)
    inner = slice(1, 11)
    assert torch.allclose(right[0, 0, inner, 0], x[0, 0, inner, 0] + 0.125, rtol=0, atol=1e-15)
    assert torch.allclose(left[0, 0, inner, 0], x[0, 0, inner, 0] - 0.125, rtol=0, atol=1e-15)
    x = _line([0.0, 1.0, 2.0, 1.0, 0.0])
    left, right = bo.reconstruct_minmod(x, 2)
    assert left[0, 0, 2, 0] == 2.0 and right[0, 0, 2, 0] == 2.0


def test_ppm_linear_constant_and_mirror():
    x = _line(1.0 - 0.3 * np.arange(16))
    left, right = bo.reconstruct_ppm(x, 2)
    inner = slice(2, 14)
    assert torch.allclose(left[0, 0, inner, 0], x[0, 0, inner, 0] + 0.15, rtol=0, atol=1e-14)
    assert torch.allclose(right[0, 0, inner, 0], x[0, 0, inner, 0] - 0.15, rtol=0, atol=1e-14)
    c = _line(np.full(8, 2.5))
    l2, r2 = bo.reconstruct_ppm(c, 2)
    assert torch.all(l2 == 2.5) and torch.all(r2 == 2.5)
    y = rng.normal(size=(5, 400))
    L, R = bo.ppm_faces(*[T(v) for v in y])
    Lm, Rm = bo.ppm_faces(*[T(v) for v in y[::-1]])
    assert torch.allclose(L, Rm, rtol=1e-13, atol=1e-13)
    assert torch.allclose(R, Lm, rtol=1e-13, atol=1e-13)


def test_ppm_monotone_data_creates_no_extrema():
    y = np.sort(rng.normal(size=(5, 2000)), axis=0)
    L, R = bo.ppm_faces(*[T(v) for v in y])
    lo, hi = T(y.min(axis=0)), T(y.max(axis=0))
    for face in (L, R):
        assert torch.all(face >= lo - 1e-12) and torch.all(face <= hi + 1e-12)


def test_weno5_beta_expanded_equals_jiang_shu():
    a, b, c = (T(rng.normal(size=1000)) for _ in range(3))
    beta = bo.weno5_beta(a, b, c) - bo.WENO_EPS * (1 + a.abs() + b.abs() + c.abs())
    js = 13 / 12 * (a - 2 * b + c) ** 2 + 0.25 * (a - 4 * b + 3 * c) ** 2
    assert torch.allclose(beta, js, rtol=1e-12, atol=1e-12)


def test_weno5_exact_on_quadratic_point_data():
    """Every substencil interpolates a quadratic exactly and the weights sum to
    one, so faces equal the quadratic at x_{i -/+ 1/2} for any weights."""
    cf = rng.normal(size=(3, 200))
    xs = np.arange(-2, 3)[:, None]
    y = cf[0] + cf[1] * xs + cf[2] * xs**2
    L, R = bo.weno5_faces(*[T(v) for v in y])
    P = lambda x: cf[0] + cf[1] * x + cf[2] * x**2
    assert np.allclose(R.numpy(), P(0.5), rtol=1e-12, atol=1e-12)
    assert np.allclose(L.numpy(), P(-0.5), rtol=1e-12, atol=1e-12)


def test_weno5_linear_data_uses_ideal_weights():
    """For linear data all three indicators are equal (b^2), so the faces are
    the ideal-weight combination, which is the 5-point Lagrange interpolant
    (3, -20, 90, 60, -5)/128 at x + 1/2: exact for quartics."""
    y = [T(3.0 + 0.7 * k) for k in range(5)]
    L, R = bo.weno5_faces(*y)
    assert abs(float(R) - (3.0 + 0.7 * 2.5)) < 1e-14
    assert abs(float(L) - (3.0 + 0.7 * 1.5)) < 1e-14
    q = rng.normal(size=5)
    vals = np.array([np.polyval(q, k) for k in range(-2, 3)])
    lag = (3 * vals[0] - 20 * vals[1] + 90 * vals[2] + 60 * vals[3] - 5 * vals[4]) / 128
    ideal = (1 / 16) * (0.375 * vals[0] - 1.25 * vals[1] + 1.875 * vals[2]) \
        + (5 / 8) * (-0.125 * vals[1] + 0.75 * vals[2] + 0.375 * vals[3]) \
        + (5 / 16) * (0.375 * vals[2] + 0.75 * vals[3] - 0.125 * vals[4])
    assert abs(lag - np.polyval(q, 0.5)) < 1e-12
    assert abs(ideal - lag) < 1e-12


# -- Riemann solvers ----------------------------------------------------------
def test_upwind_select_and_llf_identity():
    fl, fr = T(rng.normal(size=500)), T(rng.normal(size=500))
    C = T(rng.normal(size=500))
    C[:10] = 0.0
    up = bo.face_flux("upwind-flux", fl, fr, C)
    want = C * torch.where(C > 0, fl, fr)
    assert torch.equal(up, want)
    assert torch.equal(bo.riemann_upwind_state(fl, fr, C)[:10], fr[:10])  # C == 0 -> right
    llf = bo.face_flux("lax-friedrichs", fl, fr, C)
    assert torch.allclose(llf, up, rtol=0, atol=4 * 2.3e-16 * float((C.abs() * (fl.abs() + fr.abs())).max()))


# -- df/dt --------------------------------------------------------------------
def _rhs(case):
    o = oracle_solver(case)
    o._communicate_f()
    return o, o.df_dt_fvm(o.f)


def test_upwind_piecewise_constant_is_first_order_upwind():
    """pc + upwind: df/dt_i = -v (f_i - f_{i-1}) / dq for v > 0: pins the
    shift direction (a flipped shift gives downwind)."""
    c = case_2d2v(N_q=16, N_p=6, recon_q="piecewise-constant", riemann="upwind-flux",
                  source=False)
    o, L = _rhs(c)
    f = o.interior()
    Li = o.interior(L)
    v1 = o.p1_center.reshape(-1, 1, 1, 1)
    v2 = o.p2_center.reshape(-1, 1, 1, 1)
    up1 = torch.where(v1 > 0, f - torch.roll(f, 1, 2), torch.roll(f, -1, 2) - f)
    up2 = torch.where(v2 > 0, f - torch.roll(f, 1, 3), torch.roll(f, -1, 3) - f)
    want = -v1 * up1 / o.dq1 - v2 * up2 / o.dq2
    assert torch.allclose(Li, want, rtol=1e-12, atol=1e-12 * float(want.abs().max()))


@pytest.mark.parametrize("rq", ["minmod", "ppm", "weno5"])
def test_fluxes_conserve_and_constant_is_stationary(rq):
    c = case_2d2v(N_q=12, N_p=8, recon_q=rq, source=False)
    o, L = _rhs(c)
    Li = o.interior(L)
    assert abs(float(Li.sum())) <= 1e-12 * float(Li.abs().sum())
    o.f = torch.full_like(o.f, 0.7)
    o._communicate_f()
    assert float(o.interior(o.df_dt_fvm(o.f)).abs().max()) <= 1e-11


def test_bgk_conserves_moments_and_vanishes_on_maxwellian():
    c = case_2d2v(N_q=8, N_p=32)
    o = oracle_solver(c)
    rnd = torch.rand(o.f.shape, dtype=F64, generator=torch.Generator().manual_seed(0))
    f = o.f * (1 + 0.3 * rnd)
    C = bo.BGK(f, 0, o.q1_center, o.q2_center, o.p1_center, o.p2_center, o.p3_center,
               o.compute_moments, o.params)
    for name in ("density", "mom_v1_bulk", "mom_v2_bulk", "energy"):
        m = o.compute_moments(name, f=C)
        ref = o.compute_moments(name, f=f).abs().max()
        assert float(m.abs().max()) <= 1e-12 * float(ref) + 1e-14
    fmb = bo.BGK(f, 0, o.q1_center, o.q2_center, o.p1_center, o.p2_center, o.p3_center,
                 o.compute_moments, o.params, True)
    C2 = bo.BGK(fmb, 0, o.q1_center, o.q2_center, o.p1_center, o.p2_center, o.p3_center,
                o.compute_moments, o.params)
    assert float(C2.abs().max()) <= 1e-10 * float(fmb.abs().max()) / 0.01


def test_compute_moments_reference_kat():
    """bolt/lib/nonlinear/tests/test_compute_moments.py:76-144: a drifting
    Maxwellian on 16^2 x 32^3 over [-10, 10]^3; mean |error| < 1e-13."""
    N = 32
    dp = 20.0 / N
    p1, p2, p3 = bo.calculate_p_center([-10.0], [-10.0], [-10.0], N, N, N, [dp], [dp], [dp])
    _, _, _, (q1, q2) = bo.calculate_q(0, 0, 16, 16, 0, 0, 1 / 16, 1 / 16)
    rho = 1 + 0.01 * torch.sin(2 * math.pi * q1 + 4 * math.pi * q2)
    Tq = 1 + 0.01 * torch.cos(2 * math.pi * q1 + 4 * math.pi * q2)
    pb = 0.01 * torch.exp(-10 * q1**2 - 10 * q2**2)
    f = rho * (1 / (2 * math.pi * Tq)) ** 1.5 * torch.exp(-(p1 - pb) ** 2 / (2 * Tq)) \
        * torch.exp(-(p2 - pb) ** 2 / (2 * Tq)) * torch.exp(-(p3 - pb) ** 2 / (2 * Tq))
    meas = T(dp * dp * dp)
    M = bo.moments
    checks = {
        "density": rho,
        "energy": 1.5 * rho * Tq + 1.5 * rho * pb**2,
        "mom_v1_bulk": rho * pb, "mom_v2_bulk": rho * pb, "mom_v3_bulk": rho * pb,
    }
    for name, ana in checks.items():
        num = getattr(M, name)(f, p1, p2, p3, meas)
        assert float((num - ana).abs().mean()) < 1e-13, name


def _relaxation_error(stepper, dt, t_end=0.02):
    c = maxwellian_case(N_q1=1, N_q2=1, N_p1=64, N_ghost=0, p_max=10.0, p_dim=1,
                        recon_q="minmod", stepper=stepper, tau=0.01)

    def top_hat(q1, q2, p1, p2, p3, params):
        return torch.where(p1.abs() < 2.5, 1.0 + 0 * p1 * q1, 0 * p1 * q1)
    c.initialize.initialize_f = top_hat
    o = oracle_solver(c)
    f0 = o.interior().clone()
    fmb = bo.BGK(o.f, 0, o.q1_center, o.q2_center, o.p1_center, o.p2_center, o.p3_center,
                 o.compute_moments, o.params, True)
    n = int(round(t_end / dt))
    for _ in range(n):
        o.strang_timestep(dt)
    exact = fmb + (f0 - fmb) * math.exp(-n * dt / 0.01)
    return float((o.interior() - exact).abs().max())


@pytest.mark.parametrize("stepper,order", [("RK2", 2), ("SSP-RK3", 3)])
def test_time_stepper_order_on_bgk_relaxation(stepper, order):
    """The BGK relaxation of a top hat (collision_operator_test/main.py:62-97)
    with advection off is linear with a fixed Maxwellian, so f(t) is known in
    closed form; the error slope pins the RK order (test_time_integrators.py
    style)."""
    dts = [1e-3, 5e-4, 2.5e-4]
    errs = [_relaxation_error(stepper, dt) for dt in dts]
    slope = np.polyfit(np.log(dts), np.log(errs), 1)[0]
    assert abs(slope - order) < 0.2, (errs, slope)


def test_c1_runs_and_conserves_mass():
    c = case_c1(seed=0)
    o = oracle_solver(c)
    n0 = float(o.interior().sum())
    for _ in range(20):
        o.strang_timestep(c.dt)
    assert abs(float(o.interior().sum()) - n0) <= 1e-12 * n0


def test_minmod_mc_slope_on_random_line():
    """minmod.py:18-62: slope = minmod(2 (f_i - f_{i-1}), (f_{i+1} - f_{i-1}) / 2,
    2 (f_{i+1} - f_i)) on a periodic line, checked against a numpy evaluation of
    the monotonised-central limiter written from its definition."""
    v = rng.normal(size=400)
    back = v - np.roll(v, 1)
    fwd = np.roll(v, -1) - v
    cands = np.stack([2 * back, 0.5 * (back + fwd), 2 * fwd])
    agree = (np.all(cands > 0, axis=0)) | (np.all(cands < 0, axis=0))
    slope = np.where(agree, np.sign(cands[0]) * np.abs(cands).min(axis=0), 0.0)
    left, right = bo.reconstruct_minmod(_line(v), 2)
    assert np.allclose(right.reshape(-1).numpy(), v + 0.5 * slope, rtol=0, atol=1e-14)
    assert np.allclose(left.reshape(-1).numpy(), v - 0.5 * slope, rtol=0, atol=1e-14)


def test_weno5_ideal_weights_exact_on_quartics(monkeypatch):
    """With every smoothness indicator equal, the weights are the ideal ones and
    each face is the 5-point Lagrange interpolant of the stencil: exact for any
    quartic.  Pins the ideal weights and every substencil coefficient."""
    monkeypatch.setattr(bo, "weno5_beta", lambda a, b, c: torch.ones_like(a))
    cf = rng.normal(size=(5, 300))
    xs = np.arange(-2, 3)[:, None].astype(np.float64)
    y = sum(cf[k] * xs**k for k in range(5))
    L, R = bo.weno5_faces(*[T(v) for v in y])
    P = lambda x: sum(cf[k] * x**k for k in range(5))
    assert np.allclose(R.numpy(), P(0.5), rtol=0, atol=1e-12 * np.abs(cf).sum(axis=0).max())
    assert np.allclose(L.numpy(), P(-0.5), rtol=0, atol=1e-12 * np.abs(cf).sum(axis=0).max())


# -- FDTD ---------------------------------------------------------------------
def _fields_oracle(N):
    c = maxwellian_case(N_q1=N, N_q2=N, N_p1=1, N_ghost=3, p_dim=1, recon_q="minmod",
                        recon_p="minmod", source=False, fields=True)
    return oracle_solver(c).fields_solver


def _interior_yee(fs, comp):
    g1, g2 = fs.N_g1, fs.N_g2
    return fs.yee_grid_EM_fields[comp, 0, g1:g1 + fs.N_q1, g2:g2 + fs.N_q2]


def _div_b(fs):
    """Discrete divergence at the Yee nodes (i, j): B1 lives at (i+1/2, j)
    and B2 at (i, j+1/2)."""
    B1, B2 = _interior_yee(fs, 3), _interior_yee(fs, 4)
    return (B1 - torch.roll(B1, 1, 0)) / fs.dq1 + (B2 - torch.roll(B2, 1, 1)) / fs.dq2


def test_fdtd_preserves_div_b():
    """fdtd_explicit.py:84-138: the discrete curl of the Yee update has zero
    discrete divergence, so div B is constant to rounding for any E, B, J
    (test_check_maxwells_contraints.py:96 holds it to 1e-14)."""
    fs = _fields_oracle(16)
    gen = torch.Generator().manual_seed(3)
    fs.yee_grid_EM_fields = fs._bc(torch.rand(fs.yee_grid_EM_fields.shape, dtype=F64,
                                              generator=gen))
    d0 = _div_b(fs)
    sh = (1, 1) + tuple(fs.yee_grid_EM_fields.shape[2:])
    for _ in range(10):
        fs.J1, fs.J2, fs.J3 = (torch.rand(sh, dtype=F64, generator=gen) for _ in range(3))
        fs.fdtd(0.01)
    drift = float((_div_b(fs) - d0).abs().max())
    assert drift <= 1e-12 * float(d0.abs().max()), drift


def _plane_wave_error(N):
    """E3 = cos(k.q - w t), B1 = (k2/w) cos(.), B2 = -(k1/w) cos(.) solves the
    vacuum Maxwell equations (eps = mu = 1); E starts at t = 0 and B at dt/2
    as the leapfrog expects.  Returns max |E3 - exact| at t = 0.25."""
    fs = _fields_oracle(N)
    k1 = k2 = 2 * math.pi
    w = math.hypot(k1, k2)
    dt = 0.2 / N
    n = int(round(0.25 / dt))

    def cos(q1, q2, t):
        return torch.cos(k1 * q1 + k2 * q2 - w * t)[0]
    y = torch.zeros_like(fs.yee_grid_EM_fields)
    y[2] = cos(fs.q1_center, fs.q2_center, 0.0)
    y[3] = k2 / w * cos(fs.q1_center_bot, fs.q2_center_bot, dt / 2)
    # B2 sits at (i, j+1/2); the solver's q2_left_center is q2_left_bot
    # (calculate_q.py:37), so take q2 from the cell centres instead
    y[4] = -k1 / w * cos(fs.q1_left_center, fs.q2_center, dt / 2)
    fs.yee_grid_EM_fields = y
    fs.J1 = fs.J2 = fs.J3 = torch.zeros((1, 1) + tuple(y.shape[2:]), dtype=F64)
    for _ in range(n):
        fs.fdtd(dt)
    g1, g2 = fs.N_g1, fs.N_g2
    exact = cos(fs.q1_center, fs.q2_center, n * dt)[0, g1:g1 + N, g2:g2 + N]
    return float((_interior_yee(fs, 2) - exact).abs().max())


def test_fdtd_plane_wave_second_order():
    """Mode test in the style of test_fdtd_explicit.py:120-513: the error of a
    travelling plane wave falls with slope -2 under dt proportional to dq."""
    Ns = [16, 32, 64]
    errs = [_plane_wave_error(N) for N in Ns]
    slope = np.polyfit(np.log(Ns), np.log(errs), 1)[0]
    assert abs(slope + 2) < 0.25, (errs, slope)


# -- the pins must catch planted mistakes ---------------------------------------
def _mutant(fn, old, new):
    """A copy of oracle function ``fn`` with the source text ``old`` replaced
    by ``new`` (the planted mistake), compiled in the oracle's namespace."""
    import inspect
    import textwrap
    src = textwrap.dedent(inspect.getsource(fn))
    assert old in src, old
    ns = dict(vars(bo))
    exec(compile(src.replace(old, new, 1), "<mutant>", "exec"), ns)
    return ns[fn.__name__]


@pytest.mark.parametrize("target,old,new,pins", [
    ("af_shift", "shifts=ks", "shifts=[-k for k in ks]",
     ["test_af_shift_direction_and_wrap", "test_upwind_piecewise_constant_is_first_order_upwind"]),
    ("weno5_beta", "(10 / 3) * c * c", "(10 / 3) * b * c",
     ["test_weno5_beta_expanded_equals_jiang_shu"]),
    ("weno5_faces", "w2r = 5 / (8 *", "w2r = 5 / (16 *",
     ["test_weno5_ideal_weights_exact_on_quartics"]),
    ("weno5_faces", "u3l = 1.875 * y2", "u3l = 1.75 * y2",
     ["test_weno5_exact_on_quadratic_point_data"]),
    ("slope_minmod", "theta = 2\n", "theta = 1\n", ["test_minmod_mc_slope_on_random_line"]),
    ("ppm_faces", "(DQ3 - DQ2) / 6", "(DQ3 - DQ1) / 6", ["test_ppm_linear_constant_and_mirror"]),
    ("BGK_moments", "(1 / params.p_dim)", "(1 / (params.p_dim + 1))",
     ["test_bgk_conserves_moments_and_vanishes_on_maxwellian"]),
    ("f0", "(m / (2 * math.pi * k * T))\n", "(m / (math.pi * k * T))\n",
     ["test_bgk_conserves_moments_and_vanishes_on_maxwellian"]),
])
def test_pins_catch_planted_mistakes(monkeypatch, target, old, new, pins):
    monkeypatch.setattr(bo, target, _mutant(getattr(bo, target), old, new))
    g = globals()
    for name in pins:
        kwargs = {"monkeypatch": monkeypatch} if "monkeypatch" in g[name].__code__.co_varnames else {}
        with pytest.raises(AssertionError):
            g[name](**kwargs)


def test_fdtd_pins_catch_planted_mistakes(monkeypatch):
    bad_b = _mutant(bo.OracleFields.fdtd_evolve_B, "E3_minus_q2 = af_shift(E3, 0, 0, 0, 1)",
                    "E3_minus_q2 = af_shift(E3, 0, 0, 0, -1)")
    monkeypatch.setattr(bo.OracleFields, "fdtd_evolve_B", bad_b)
    with pytest.raises(AssertionError):
        test_fdtd_preserves_div_b()
    monkeypatch.undo()
    bad_e = _mutant(bo.OracleFields.fdtd_evolve_E, "- (B1_plus_q2 - B1) / self.dq2",
                    "+ (B1_plus_q2 - B1) / self.dq2")
    monkeypatch.setattr(bo.OracleFields, "fdtd_evolve_E", bad_e)
    with pytest.raises(AssertionError):
        test_fdtd_plane_wave_second_order()


# -- electrostatic FFT Poisson --------------------------------------------------
def _single_mode_error(k1, k2, N1=16, N2=12, eps=1.0):
    """fft.py:9-78 on rho = 3 + A cos(2 pi (k1 x + k2 y)) at cell centres: the
    mean is removed and the field is spectrally exact,
    E = -grad phi with lap phi = -rho, i.e.
    E = A (k1, k2) sin(.) / (2 pi |k|^2 eps)."""
    L1, L2 = 1.0, 2.0
    x = (np.arange(N1) + 0.5) * L1 / N1
    y = (np.arange(N2) + 0.5) * L2 / N2
    X, Y = np.meshgrid(x, y, indexing="ij")
    A = 0.7
    ph = 2 * np.pi * (k1 * X / L1 + k2 * Y / L2)
    rho = 3.0 + A * np.cos(ph)
    E1, E2 = bo.fft_poisson(T(rho), L1 / N1, L2 / N2, eps)
    kk1, kk2 = k1 / L1, k2 / L2
    ksq = kk1 ** 2 + kk2 ** 2
    w1 = A * kk1 * np.sin(ph) / (2 * np.pi * ksq * eps)
    w2 = A * kk2 * np.sin(ph) / (2 * np.pi * ksq * eps)
    return max(np.abs(E1.numpy() - w1).max(), np.abs(E2.numpy() - w2).max())


@pytest.mark.parametrize("k1,k2", [(1, 0), (0, 2), (2, 1), (3, -2)])
def test_fft_poisson_single_mode_closed_form(k1, k2):
    assert _single_mode_error(k1, k2) <= 1e-13
    assert _single_mode_error(k1, k2, eps=2.5) <= 1e-13


def test_fft_poisson_catches_planted_mistakes(monkeypatch):
    good = bo.fft_poisson
    monkeypatch.setattr(bo, "fft_poisson", _mutant(good, "E1 = torch.fft.ifft2(-1j", "E1 = torch.fft.ifft2(1j"))
    assert _single_mode_error(2, 1) > 1e-3
    monkeypatch.setattr(bo, "fft_poisson", _mutant(good, "4 * math.pi ** 2", "2 * math.pi ** 2"))
    assert _single_mode_error(2, 1) > 1e-3
