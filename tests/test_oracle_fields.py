"""Pins of the oracle's electrodynamic path (CPU): the Yee <-> cell-centre
averaging, the v x B and E terms of C_p driving the p-advection, and the time
levels the RK stages see, each against what the mathematics fixes, with
planted mistakes that must fail.

- fields.py:497-557 (Yee <-> cell centre): an average of two (four) staggered
  values of a Fourier mode sin(2 pi (k1 x + k2 y) + phi) is the mode at the
  midpoint times cos(pi k h) per averaged axis, exactly;
- advection_terms.py:147-188 with df_dt_fvm.py:251-323 (C_p and the p-fluxes):
  for a q-uniform drifting Maxwellian with a piecewise-constant p-reconstruction
  the upwind face states along each p-axis sum to the cell values (the face
  velocity is uniform along its own axis), so the bulk velocity obeys
  du/dt = (q/m)(E + u x B) exactly in semi-discrete form: it rotates at the
  cyclotron frequency qB/m in a uniform B3 and drifts linearly in a uniform E1;
- fields.py:568-690 with fvm_operator.py:22-33 (time levels): the k = 0
  Langmuir oscillation of a uniform plasma is the linear system
  d(n u)/dt = (q/m) n E, dE/dt = -q n u / eps, so each stepper's
  (momentum, field) sequence is a fixed linear recurrence, and its frequency
  tends to omega_p = sqrt(n q^2 / (m eps)) as dt -> 0; a vacuum wave seen by
  the stages of SSP-RK3 (charge 0) sits at t^n, t^{n+1} and t^{n+1/2} up to
  O(dt^2).
"""
import math
import types

import numpy as np
import pytest
import torch

from oracle import bolt_oracle as bo
from tests.cases import maxwellian_case

F64 = torch.float64


# -- Yee <-> cell centre --------------------------------------------------------
def _yee_oracle(N1=12, N2=10):
    c = maxwellian_case(N_q1=N1, N_q2=N2, N_p1=1, N_ghost=3, p_dim=1, recon_q="minmod",
                        recon_p="minmod", source=False, fields=True)
    c.domain.q2_end = 1.3  # dq1 != dq2
    from tests.cases import oracle_solver
    return oracle_solver(c).fields_solver


def _mode(fs, k1, k2, phi, dx1, dx2):
    """sin(2 pi (k1 x + k2 y) + phi) at x = (i + dx1) h1, y = (j + dx2) h2 over
    the padded grid (ghosts included: the mode is periodic), as (1, Q1g, Q2g)."""
    g1, g2 = fs.N_g1, fs.N_g2
    i = torch.arange(-g1, fs.N_q1 + g1, dtype=F64).reshape(-1, 1)
    j = torch.arange(-g2, fs.N_q2 + g2, dtype=F64).reshape(1, -1)
    L1, L2 = fs.N_q1 * fs.dq1, fs.N_q2 * fs.dq2
    x, y = (i + dx1) * fs.dq1, (j + dx2) * fs.dq2
    return torch.sin(2 * math.pi * (k1 * x / L1 + k2 * y / L2) + phi).unsqueeze(0)


# Yee offsets (i + dx1, j + dx2) of E1, E2, E3, B1, B2, B3 (fields.py:527-557)
YEE_AT = [(0.0, 0.5), (0.5, 0.0), (0.5, 0.5), (0.5, 0.0), (0.0, 0.5), (0.0, 0.0)]


def _averaging_factor(fs, k1, k2, at):
    c1 = math.cos(math.pi * k1 / fs.N_q1)
    c2 = math.cos(math.pi * k2 / fs.N_q2)
    return (c1 if at[0] != 0.5 else 1.0) * (c2 if at[1] != 0.5 else 1.0)


def _yee_cell_errors(k1=2, k2=-3, phi=0.4):
    fs = _yee_oracle()
    g1, g2 = fs.N_g1, fs.N_g2
    inner = (slice(None), slice(g1, g1 + fs.N_q1), slice(g2, g2 + fs.N_q2))
    # Yee -> cell centre
    y = torch.stack([_mode(fs, k1, k2, phi + c, *YEE_AT[c]) for c in range(6)])
    fs.yee_grid_EM_fields = y.clone()
    fs.yee_grid_to_cell_centered_grid()
    err_to_cc = 0.0
    for c in range(6):
        want = _averaging_factor(fs, k1, k2, YEE_AT[c]) * _mode(fs, k1, k2, phi + c, 0.5, 0.5)
        got = fs.cell_centered_EM_fields[c]
        err_to_cc = max(err_to_cc, float((got[inner] - want[inner]).abs().max()))
    # cell centre -> Yee
    fs.cell_centered_EM_fields = torch.stack(
        [_mode(fs, k1, k2, phi + c, 0.5, 0.5) for c in range(6)])
    fs.yee_grid_EM_fields = torch.zeros_like(y)
    fs.cell_centered_grid_to_yee_grid()
    err_to_yee = 0.0
    for c in range(6):
        want = _averaging_factor(fs, k1, k2, YEE_AT[c]) * _mode(fs, k1, k2, phi + c, *YEE_AT[c])
        got = fs.yee_grid_EM_fields[c]
        err_to_yee = max(err_to_yee, float((got[inner] - want[inner]).abs().max()))
    return err_to_cc, err_to_yee


@pytest.mark.parametrize("k1,k2", [(1, 0), (0, 1), (2, -3), (3, 2)])
def test_yee_cell_averaging_fourier_exact(k1, k2):
    a, b = _yee_cell_errors(k1, k2)
    assert a <= 1e-14 and b <= 1e-14, (a, b)


def test_yee_cell_pins_catch_planted_mistakes(monkeypatch):
    from tests.test_oracle_pins import _mutant
    for fn, old, new in [
        (bo.OracleFields.yee_grid_to_cell_centered_grid,
         "(E1 + af_shift(E1, 0, 0, -1, 0))", "(E1 + af_shift(E1, 0, 0, 1, 0))"),
        (bo.OracleFields.yee_grid_to_cell_centered_grid,
         "(B1 + af_shift(B1, 0, 0, 0, -1))", "(B1 + af_shift(B1, 0, 0, -1, 0))"),
        (bo.OracleFields.cell_centered_grid_to_yee_grid,
         "+ af_shift(B3, 0, 0, 1, 1)", "+ af_shift(B3, 0, 0, -1, 1)"),
        (bo.OracleFields.cell_centered_grid_to_yee_grid,
         "(E2 + af_shift(E2, 0, 0, 0, 1))", "(E2 + af_shift(E2, 0, 0, 1, 0))"),
    ]:
        monkeypatch.setattr(bo.OracleFields, fn.__name__, _mutant(fn, old, new))
        a, b = _yee_cell_errors(2, -3)
        assert max(a, b) > 1e-3, (fn.__name__, old)
        monkeypatch.undo()


# -- C_p on a q-uniform drifting Maxwellian --------------------------------------
def uniform_plasma_case(*, stepper, u0=(0.0, 0.0), E0=(0.0, 0.0, 0.0), B0=(0.0, 0.0, 0.0),
                        eps=1.0, charge=-1.0, mass=1.0, N_p=32, p_max=8.0, p_dim=2,
                        recon_p="piecewise-constant"):
    """A drifting Maxwellian (T = 1, n = 1) uniform on a periodic 4 x 4 q-grid
    with uniform E and B, fields on (FDTD), BGK off (the user modules)."""
    c = maxwellian_case(N_q1=4, N_q2=4, N_p1=N_p, N_p2=N_p if p_dim >= 2 else 1,
                        N_p3=N_p if p_dim >= 3 else 1, N_ghost=3, p_max=p_max, p_dim=p_dim,
                        recon_q="minmod", recon_p=recon_p, riemann="upwind-flux",
                        stepper=stepper, source=False, fields=True, charge=charge, mass=mass)
    p = c.params
    p.eps, p.mu, p.u0, p.E0, p.B0v = eps, 1.0, tuple(u0), tuple(E0), tuple(B0)

    def initialize_f(q1, q2, p1, p2, p3, params):
        m = params.mass
        e = (p1 - params.u0[0]) ** 2
        if params.p_dim >= 2:
            e = e + (p2 - params.u0[1]) ** 2
        if params.p_dim >= 3:
            e = e + p3 ** 2
        return (q1 * 0 + 1) * (m / (2 * math.pi)) ** (params.p_dim / 2) * torch.exp(-m * e / 2)

    def initialize_E(q1, q2, params):
        z = 0.0 * q1
        return tuple(z + v for v in params.E0)

    def initialize_B(q1, q2, params):
        z = 0.0 * q1
        return tuple(z + v for v in params.B0v)

    c.initialize.initialize_f = initialize_f
    c.initialize.initialize_E = initialize_E
    c.initialize.initialize_B = initialize_B
    return c


def _uniform_plasma(*, advection=None, **kw):
    c = uniform_plasma_case(**kw)
    sysm = bo.make_system(c.domain, c.boundary_conditions, c.params, c.initialize,
                          advection or bo.advection_terms, bo.BGK, bo.moments)
    return bo.OracleSolver(sysm)


def _bulk(o):
    """(n, n u1, n u2, E1) at one interior q-cell (the state is q-uniform)."""
    g = o.N_g1
    n = float(o.compute_moments("density")[0, 0, g, g])
    m1 = float(o.compute_moments("mom_v1_bulk")[0, 0, g, g])
    m2 = float(o.compute_moments("mom_v2_bulk")[0, 0, g, g])
    E1 = float(o.fields_solver.yee_grid_EM_fields[0, 0, g, g])
    return n, m1, m2, E1


@pytest.mark.parametrize("stepper", ["RK2", "SSP-RK3"])
def test_cp_gyration_in_uniform_b3(stepper):
    """du/dt = (q/m) u x B: with B = (0, 0, B0) the bulk velocity rotates,
    u1 + i u2 = u0 exp(-i w t), w = q B0 / m.  The stepper applies its
    stability polynomial R(z) (RK2: 1 + z + z^2/2; SSP-RK3: + z^3/6) to
    z = -i w dt per step, so u1 + i u2 = u0 R(z)^N exactly; both tend to the
    continuous rotation.  E stays ~0 (eps = 1e12: the current of the drifting
    plasma cannot move it)."""
    qm, B0, u0, dt, n_steps = -1.0, 0.5, 0.6, 0.02, 100
    # p in [-16, 16]: the first-order upwinding in p spreads f, and the
    # identity needs f ~ 0 where the p-axes wrap around
    o = _uniform_plasma(stepper=stepper, u0=(u0, 0.0), B0=(0.0, 0.0, B0), eps=1e12,
                        N_p=64, p_max=16.0)
    for _ in range(n_steps):
        o.strang_timestep(dt)
    n, m1, m2, _ = _bulk(o)
    w, t = qm * B0, n_steps * dt
    z = -1j * w * dt
    R = 1 + z + z * z / 2 + (z ** 3 / 6 if stepper == "SSP-RK3" else 0)
    want = u0 * R ** n_steps
    err = abs(complex(m1 / n, m2 / n) - want)
    assert err <= 1e-11 * u0, (stepper, m1 / n, m2 / n, want, err)
    assert abs(want - u0 * complex(math.cos(w * t), -math.sin(w * t))) <= 5e-5 * u0


@pytest.mark.parametrize("stepper", ["RK2", "SSP-RK3"])
def test_cp_acceleration_in_uniform_e1(stepper):
    """du1/dt = (q/m) E1 with E1 uniform (eps = 1e12 keeps it fixed): u1 is
    linear in t, which both steppers integrate exactly."""
    qm, E1, u0, dt, n_steps = -1.0, 0.3, 0.2, 0.02, 50
    o = _uniform_plasma(stepper=stepper, u0=(u0, 0.0), E0=(E1, 0.0, 0.0), eps=1e12)
    for _ in range(n_steps):
        o.strang_timestep(dt)
    n, m1, m2, _ = _bulk(o)
    want = u0 + qm * E1 * n_steps * dt
    assert abs(m1 / n - want) <= 1e-9 and abs(m2 / n) <= 1e-12, (m1 / n, want, m2 / n)


def test_cp_pins_catch_planted_mistakes(monkeypatch):
    from tests.test_oracle_pins import _mutant
    for old, new, pin in [
        ("(E1 + v2 * B3 - v3 * B2)", "(E1 - v2 * B3 - v3 * B2)", test_cp_gyration_in_uniform_b3),
        ("(E2 + v3 * B1 - v1 * B3)", "(E2 + v3 * B1 + v1 * B3)", test_cp_gyration_in_uniform_b3),
        ("(E1 + v2 * B3 - v3 * B2)", "(-E1 + v2 * B3 - v3 * B2)",
         test_cp_acceleration_in_uniform_e1),
    ]:
        monkeypatch.setattr(bo, "C_p", _mutant(bo.C_p, old, new))
        monkeypatch.setattr(bo._OraclePlugins, "C_p", staticmethod(bo.C_p))
        with pytest.raises(AssertionError):
            pin("SSP-RK3")
        monkeypatch.undo()
    # the p-flux difference applied with the wrong sign
    bad = _mutant(bo.OracleSolver.df_dt_fvm, "df_dt = df_dt - (d1 + d2 + d3)",
                  "df_dt = df_dt + (d1 + d2 + d3)")
    monkeypatch.setattr(bo.OracleSolver, "df_dt_fvm", bad)
    with pytest.raises(AssertionError):
        test_cp_acceleration_in_uniform_e1("RK2")


# -- k = 0 Langmuir oscillation: time levels of the field coupling -------------
def _langmuir_trajectory(stepper, dt, n_steps, u0=0.05):
    o = _uniform_plasma(stepper=stepper, u0=(u0,), eps=1.0, N_p=48, p_dim=1)
    traj = [_bulk(o)]
    for _ in range(n_steps):
        o.strang_timestep(dt)
        traj.append(_bulk(o))
    return o, traj


def langmuir_recurrence(stepper, dt, n_steps, n, m, E, qm=-1.0, q=-1.0, eps=1.0):
    """(momentum n u1, E1) after each step of the scheme's own definition for
    the k = 0 mode (a = q/m, J = q n u1, the leapfrog E += -dt J / eps):

    RK2 with the reference's field coupling (df_dt_fvm.py:99-104 with
    fields.py:568-690): stage 1 runs no FDTD and sees X, the cell-centred E1;
    stage 2 runs the leapfrog with J of the stage-1 output and then sees X,
    which fields.py:493-494 binds to the same array as the time levels, so the
    final write of fields.py:642-643 leaves X <- (X + E^{n+1}) / 2.

    SSP-RK3 with the once-per-step coupling: stage 1 sees E^n, the leapfrog
    uses J^{n+1/2} = (J(u^n) + J(u1)) / 2, stage 2 sees E^{n+1}, stage 3
    (E^n + E^{n+1}) / 2."""
    out = [(m, E)]
    X = E
    for _ in range(n_steps):
        if stepper == "RK2":
            mh = m + 0.5 * dt * qm * n * X
            E = E - dt * (q * mh) / eps
            X = 0.5 * (X + E)
            m = m + dt * qm * n * X
        else:
            m1 = m + dt * qm * n * E
            En = E - dt * (0.5 * (q * m + q * m1)) / eps
            m2 = 0.75 * m + 0.25 * (m1 + dt * qm * n * En)
            m = m / 3 + (2 / 3) * (m2 + dt * qm * n * (0.5 * (E + En)))
            E = En
        out.append((m, E))
    return out


def _frequency(ts, us):
    """Angular frequency from the first two downward/upward zero crossings of
    u(t) (linear interpolation; u ~ cos near a crossing, so O(dt^3))."""
    zs = []
    for k in range(len(us) - 1):
        if us[k] == 0.0 or us[k] * us[k + 1] < 0:
            zs.append(ts[k] + (ts[k + 1] - ts[k]) * us[k] / (us[k] - us[k + 1]))
    assert len(zs) >= 2, "run shorter than a period"
    return math.pi / (zs[1] - zs[0])


@pytest.mark.parametrize("stepper", ["RK2", "SSP-RK3"])
def test_langmuir_k0_follows_the_scheme_recurrence(stepper):
    dt, n_steps = 0.05, 40
    _, traj = _langmuir_trajectory(stepper, dt, n_steps)
    n, m0, _, E0 = traj[0]
    rec = langmuir_recurrence(stepper, dt, n_steps, n, m0, E0)
    scale_m = max(abs(t[1]) for t in traj)
    scale_E = max(abs(t[3]) for t in traj)
    for (nn, m, _, E), (mr, Er) in zip(traj, rec):
        assert abs(m - mr) <= 1e-12 * scale_m and abs(E - Er) <= 1e-12 * scale_E, (m, mr, E, Er)


def _langmuir_measure(stepper, dt):
    """(relative frequency error against omega_p, growth rate of the
    amplitude per unit time) of the oracle's k = 0 oscillation over t = 10
    (about 1.6 periods).  The invariant of d(n u)/dt = (q/m) n E,
    dE/dt = -q n u / eps is (q/eps) (n u)^2 + (q/m) n E^2, i.e. with
    q = -1, m = eps = 1 the squared amplitude (n u)^2 + n E^2."""
    n_steps = int(math.ceil(10.0 / dt))
    _, traj = _langmuir_trajectory(stepper, dt, n_steps)
    n = traj[0][0]
    ts = [k * dt for k in range(len(traj))]
    w = _frequency(ts, [t[1] / n for t in traj])
    omega_p = math.sqrt(n * 1.0 / (1.0 * 1.0))  # n q^2 / (m eps)
    amp2 = [t[1] ** 2 + n * t[3] ** 2 for t in traj]
    # skip the first steps (the RK2 recurrence's third eigenvalue, 1/2, decays)
    k0 = 10
    rate = 0.5 * math.log(amp2[-1] / amp2[k0]) / (ts[-1] - ts[k0])
    return abs(w - omega_p) / omega_p, rate


@pytest.mark.parametrize("stepper", ["RK2", "SSP-RK3"])
def test_langmuir_k0_frequency_converges_to_plasma_frequency(stepper):
    """The measured frequency tends to omega_p at second order as dt halves,
    for the reference's RK2 coupling and for SSP-RK3 with its once-per-step
    coupling.  The amplitude: SSP-RK3's coupling is neutral on this mode
    (its recurrence has |eigenvalue| = 1); the reference's RK2 coupling grows
    it at a rate that falls at first order (about omega_p^2 dt / 4: its stage
    fields are the running average that fields.py:493-494 and :642-643 make
    of E, which lags E by about a step)."""
    dts = [0.1, 0.05, 0.025]
    res = [_langmuir_measure(stepper, dt) for dt in dts]
    errs, rates = [r[0] for r in res], [r[1] for r in res]
    slope = np.polyfit(np.log(dts), np.log(errs), 1)[0]
    assert errs[-1] <= 1e-3 and abs(slope - 2) <= 0.3, (errs, slope)
    if stepper == "SSP-RK3":
        # (the continuous invariant differs from the discrete one by O(dt^2),
        # which the endpoints' phases show as a small apparent rate)
        assert max(abs(r) for r in rates) <= 2e-4, rates
    else:
        rslope = np.polyfit(np.log(dts), np.log(rates), 1)[0]
        assert min(rates) > 0 and abs(rslope - 1) <= 0.2, (rates, rslope)
        assert all(abs(r / (0.25 * dt) - 1) <= 0.1 for r, dt in zip(rates, dts)), rates


def test_langmuir_pins_catch_planted_mistakes(monkeypatch):
    # the reference's per-C_p toggle applied to SSP-RK3 (FDTD on 1, 2, 1, 2, ...
    # stages: the fields advance 1.5 dt per step)
    orig = bo.OracleFields.__init__

    def toggled(self, solver, rho):
        orig(self, solver, rho)
        self.coupling = "toggle"
        self.cell_centered_EM_fields_at_n = self.cell_centered_EM_fields
        self.cell_centered_EM_fields_at_n_plus_half = self.cell_centered_EM_fields
    monkeypatch.setattr(bo.OracleFields, "__init__", toggled)
    assert _langmuir_measure("SSP-RK3", 0.05)[0] > 0.1
    with pytest.raises(AssertionError):
        test_langmuir_k0_follows_the_scheme_recurrence("SSP-RK3")
    monkeypatch.undo()
    # stage 3 of SSP-RK3 fed the fields at n+1 instead of n+1/2
    from tests.test_oracle_pins import _mutant
    bad = _mutant(bo.OracleSolver.timestep_fvm_ssprk3,
                  "fs.stage_fields = fs.cell_centered_EM_fields_at_n_plus_half",
                  "fs.stage_fields = fs.cell_centered_EM_fields_at_n")
    monkeypatch.setattr(bo.OracleSolver, "timestep_fvm_ssprk3", bad)
    with pytest.raises(AssertionError):
        test_langmuir_k0_follows_the_scheme_recurrence("SSP-RK3")
    with pytest.raises(AssertionError):
        test_langmuir_k0_frequency_converges_to_plasma_frequency("SSP-RK3")


# -- stage time levels on a vacuum wave (SSP-RK3) ------------------------------
def _recording_advection(log):
    def C_p(t, q1, q2, v1, v2, v3, fields_solver, params):
        fields = fields_solver.get_fields()
        log.append(torch.cat(fields).clone())
        z = 0.0 * fields[0] * v1
        return z, z, z
    return types.SimpleNamespace(C_q=bo.C_q, A_q=bo.A_q, C_p=C_p, A_p=C_p)


def _stage_level_error(N, levels=(0.0, 1.0, 0.5)):
    """Charge 0 (no current, f frozen), a vacuum plane wave E3 = cos(k.q - w t),
    B1 = (k2/w) cos, B2 = -(k1/w) cos with B started at dt/2 as the leapfrog
    expects; step 1's three stages (3 C_p calls each) must see the
    cell-centred wave at t^1 + levels[s] dt.  Returns the largest deviation."""
    log = []
    c = maxwellian_case(N_q1=N, N_q2=N, N_p1=4, N_ghost=3, p_max=2.0, p_dim=1,
                        recon_q="minmod", recon_p="piecewise-constant",
                        riemann="upwind-flux", stepper="SSP-RK3", source=False, fields=True,
                        charge=0.0)
    c.initialize = types.SimpleNamespace(**vars(c.initialize))
    sysm = bo.make_system(c.domain, c.boundary_conditions, c.params, c.initialize,
                          _recording_advection(log), bo.BGK, bo.moments)
    o = bo.OracleSolver(sysm)
    fs = o.fields_solver
    k1, k2 = 2 * math.pi, 2 * math.pi
    w = math.hypot(k1, k2)
    dt = 0.2 / N

    def wave(dx1, dx2, t):
        g = fs.N_g1
        i = torch.arange(-g, N + g, dtype=F64).reshape(-1, 1)
        j = torch.arange(-g, N + g, dtype=F64).reshape(1, -1)
        return torch.cos(k1 * (i + dx1) / N + k2 * (j + dx2) / N - w * t).unsqueeze(0)
    y = torch.zeros_like(fs.yee_grid_EM_fields)
    y[2] = wave(0.5, 0.5, 0.0)
    y[3] = k2 / w * wave(0.5, 0.0, dt / 2)
    y[4] = -k1 / w * wave(0.0, 0.5, dt / 2)
    fs.yee_grid_EM_fields = y
    fs.yee_grid_to_cell_centered_grid()
    fs.cell_centered_EM_fields_at_n = fs.cell_centered_EM_fields.clone()
    fs.cell_centered_EM_fields_at_n_plus_half = fs.cell_centered_EM_fields.clone()
    o.strang_timestep(dt)
    log.clear()
    o.strang_timestep(dt)
    assert len(log) == 9
    c1, c2 = math.cos(math.pi / N), math.cos(math.pi / N)
    g = fs.N_g1
    inner = (slice(g, g + N), slice(g, g + N))
    err = 0.0
    for s, lv in enumerate(levels):
        t = (1 + lv) * dt
        want = {2: wave(0.5, 0.5, t)[0], 3: c2 * k2 / w * wave(0.5, 0.5, t)[0],
                4: -c1 * k1 / w * wave(0.5, 0.5, t)[0]}
        for call in range(3):
            got = log[3 * s + call][:, 0]  # (6, Q1g, Q2g)
            for comp, val in want.items():
                err = max(err, float((got[comp][inner] - val[inner]).abs().max()))
            for comp in (0, 1, 5):
                err = max(err, float(got[comp][inner].abs().max()))
    return err


def test_rk3_stages_see_fields_at_their_time_levels():
    """SSP-RK3's stages sit at t^n, t^{n+1}, t^{n+1/2}: the fields each one
    sees deviate from the exact wave there by O(dt^2) (dt ~ dq), so the error
    falls with slope -2; a stage fed the wrong time level would be off by
    O(dt)."""
    Ns = [16, 32, 64]
    errs = [_stage_level_error(N) for N in Ns]
    slope = np.polyfit(np.log(Ns), np.log(errs), 1)[0]
    assert abs(slope + 2) <= 0.3, (errs, slope)


def test_stage_level_pin_catches_planted_mistakes(monkeypatch):
    from tests.test_oracle_pins import _mutant
    for old, new in [
        ("fs.stage_fields = fs.cell_centered_EM_fields_at_n_plus_half",
         "fs.stage_fields = fs.cell_centered_EM_fields_at_n"),
        # (the method's source is dedented by one level in the mutant)
        ("        fs.stage_fields = fs.cell_centered_EM_fields_at_n\n"
         "    u2 = ",
         "        fs.stage_fields = fs.cell_centered_EM_fields_at_n_plus_half\n"
         "    u2 = "),
    ]:
        bad = _mutant(bo.OracleSolver.timestep_fvm_ssprk3, old, new)
        monkeypatch.setattr(bo.OracleSolver, "timestep_fvm_ssprk3", bad)
        with pytest.raises(AssertionError):
            test_rk3_stages_see_fields_at_their_time_levels()
        monkeypatch.undo()
    # B at n taken as B^{n+1/2} (no average of the half steps)
    bad = _mutant(bo.OracleFields.evolve_once_per_step,
                  "0.5 * (B_n_plus_half + B_n_plus_3half)", "B_n_plus_3half")
    monkeypatch.setattr(bo.OracleFields, "evolve_once_per_step", bad)
    with pytest.raises(AssertionError):
        test_rk3_stages_see_fields_at_their_time_levels()
