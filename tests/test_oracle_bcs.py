"""Pins of the oracle's non-periodic q boundaries (boundaries.py:147-577 with
the 'ghosted' DMDA of communicate.py): invariants and a from-the-definition
scheme, each failing for a plausible mistake planted in a copy of the oracle
(no velocity flip at a mirror wall, a shifted mirror image, the inflow select
on the wrong sign)."""
import math

import numpy as np
import pytest
import torch

from oracle import bolt_oracle as bo
from tests.cases import case_bc, drifting_maxwellian, oracle_solver

F64 = torch.float64


def _mass(o):
    return float(o.interior().sum())


def _run(c, steps):
    o = oracle_solver(c)
    m0 = _mass(o)
    for _ in range(steps):
        o.strang_timestep(c.dt)
    return o, m0


# -- mirror walls ---------------------------------------------------------------
@pytest.mark.parametrize("recon", ["piecewise-constant", "minmod", "ppm"])
def test_mirror_walls_conserve_mass(recon):
    """With mirror ghosts the flux pair (+v, -v) through a wall cancels (the
    reconstruction is mirror-symmetric and the p-grid symmetric), so the total
    of f is conserved to rounding; a drifting state makes every wall carry
    flux.  The reference's WENO5 is not mirror-symmetric (weno5.py:58-71 uses
    the first substencil's indicator formula for all three), so it is not
    pinned this way."""
    c = case_bc(left="mirror", right="mirror", bottom="mirror", top="mirror",
                recon_q=recon, drift=(0.9, -0.6))
    o, m0 = _run(c, 4)
    assert abs(_mass(o) - m0) <= 1e-13 * abs(m0)
    # the walls are active: the state is no longer its own periodic image
    assert float((o.interior() - oracle_solver(c).interior()).abs().max()) > 1e-6


@pytest.mark.parametrize("recon", ["minmod", "ppm"])
def test_mirror_walls_conserve_mass_with_electrostatic_fields(recon):
    """Electrostatic fields move f in p only, and the p-fluxes telescope on
    the periodic p-grid, so mirror walls still conserve the total of f to
    rounding; the fields are active (the p-advection changes the state)."""
    kw = dict(left="mirror", right="mirror", bottom="mirror", top="mirror",
              recon_q=recon, drift=(0.9, -0.6), bump=0.5)
    c = case_bc(electrostatic=True, **kw)
    o, m0 = _run(c, 4)
    assert abs(_mass(o) - m0) <= 1e-13 * abs(m0)
    o0, _ = _run(case_bc(**kw), 4)
    assert float((o.interior() - o0.interior()).abs().max()) > 1e-8


def test_mirror_keeps_mirror_symmetric_state_symmetric():
    """f(q1, p1) = f(1 - q1, -p1) in a box with mirror walls left and right
    stays mirror-symmetric: the two walls are handled alike."""
    c = case_bc(left="mirror", right="mirror", N_q2=1, p_dim=1, N_p1=8, recon_q="ppm",
                drift=(0.0, 0.0), stepper="SSP-RK3")

    def init(q1, q2, p1, p2, p3, params):
        n = 1.0 + 0.3 * torch.exp(-(q1 - 0.3) ** 2 / 0.02) + 0.3 * torch.exp(-(q1 - 0.7) ** 2 / 0.02)
        skew = 1.0 + 0.2 * torch.tanh(4 * (q1 - 0.5)) * torch.tanh(p1)
        return n * skew * torch.exp(-p1 ** 2 / 2) / math.sqrt(2 * math.pi)

    c.initialize.initialize_f = init
    o, _ = _run(c, 5)
    f = o.interior()[:, 0, :, 0]                        # (p1, q1)
    g = torch.flip(f, (0, 1))                           # p1 -> -p1, q1 -> 1 - q1
    assert float((f - g).abs().max()) <= 1e-14 * float(f.abs().max())
    # the state moved: the symmetry is not that of the initial data alone
    assert float((f - oracle_solver(c).interior()[:, 0, :, 0]).abs().max()) > 1e-6


# -- dirichlet inflow ---------------------------------------------------------
def _upwind_rk2_by_definition(f, v, dq, dt, g_left, g_right, steps):
    """First-order upwind FVM for df/dt + v df/dq = 0 on N cells with inflow
    values g_left (used where v > 0) and g_right (v < 0), RK2 midpoint, in
    plain numpy loops."""
    def L(u):
        N, P = u.shape
        out = np.zeros_like(u)
        for j in range(P):
            for i in range(N):
                # flux through the left face i-1/2 and the right face i+1/2
                def face(k):  # face between cells k-1 and k
                    if v[j] > 0:
                        return v[j] * (u[k - 1, j] if k >= 1 else g_left[j])
                    return v[j] * (u[k, j] if k <= N - 1 else g_right[j])
                out[i, j] = -(face(i + 1) - face(i)) / dq
        return out
    for _ in range(steps):
        half = f + 0.5 * dt * L(f)
        f = f + dt * L(half)
    return f


def test_dirichlet_piecewise_constant_is_upwind_with_inflow_values():
    c = case_bc(left="dirichlet", right="dirichlet", N_q1=10, N_q2=1, p_dim=1, N_p1=6,
                recon_q="piecewise-constant", inflow=(1.3, 0.4, 0.0), drift=(-0.5, 0.0))
    o = oracle_solver(c)
    f0 = o.interior()[:, 0, :, 0].T.numpy().copy()      # (q1, p1)
    v = o.p1_center.reshape(-1).numpy()
    g = (1.3 * np.exp(-(v - 0.4) ** 2 / 2) / math.sqrt(2 * math.pi))
    for _ in range(3):
        o.strang_timestep(c.dt)
    want = _upwind_rk2_by_definition(f0, v, 1.0 / 10, c.dt, g, g, 3)
    got = o.interior()[:, 0, :, 0].T.numpy()
    assert np.abs(got - want).max() <= 1e-14 * np.abs(want).max()


@pytest.mark.parametrize("recon", ["piecewise-constant", "minmod"])
def test_uniform_state_equal_to_inflow_is_stationary(recon):
    """A uniform state equal to the inflow Maxwellian on both sides: outflow
    ghosts are zero, but the limited slopes at the boundary cells vanish, so
    nothing moves."""
    c = case_bc(left="dirichlet", right="dirichlet", bottom="dirichlet", top="dirichlet",
                recon_q=recon, inflow=(1.1, 0.3, -0.2), drift=(0.3, -0.2), bump=0.0)
    c.initialize.initialize_f = lambda q1, q2, p1, p2, p3, params: 1.1 * (
        2 * math.pi) ** -1 * torch.exp(-((p1 - 0.3) ** 2 + (p2 + 0.2) ** 2) / 2) * (q1 * 0 + 1)
    o, _ = _run(c, 3)
    f = o.interior()
    ref = oracle_solver(c).interior()
    assert float((f - ref).abs().max()) <= 1e-15 * float(ref.abs().max())


def test_mirror_plus_dirichlet_takes_inflow_and_mirrors_outflow():
    """'mirror+dirichlet' mirrors first, then overwrites the inflow half: the
    left ghost zone equals the user's values where C_q1 > 0 and the mirror
    image (p1 reversed) elsewhere."""
    c = case_bc(left="mirror+dirichlet", right="periodic", N_q1=8, N_q2=1, p_dim=1,
                N_p1=6, inflow=(1.4, 0.2, 0.0))
    c.boundary_conditions.in_q1_right = "mirror+dirichlet"
    o = oracle_solver(c)
    o._communicate_f()
    g = o.N_ghost
    v = o.p1_center.reshape(-1)
    user = drifting_maxwellian(v, 0 * v, 0 * v, 1, 1.4, 0.2, 0.0)
    interior = o.f[:, 0, g:g + 8, 0]
    for k in range(g):                                  # ghost cell -1-k mirrors cell k
        ghost = o.f[:, 0, g - 1 - k, 0]
        mirror = torch.flip(interior[:, k], (0,))
        want = torch.where(v > 0, user, mirror)
        assert torch.equal(ghost, want)


# -- planted mistakes ---------------------------------------------------------
@pytest.mark.parametrize("old,new,pin", [
    # no velocity reversal at the left wall
    ("torch.flip(self._convert_to_p_expanded(f), (0,)))[:, :, :g]",
     "self._convert_to_p_expanded(f))[:, :, :g]", "mass"),
    # the mirror image shifted by one cell at the bottom wall
    ("f[:, :, :, :g] = torch.flip(f[:, :, :, g:2 * g], (3,))",
     "f[:, :, :, :g] = torch.flip(f[:, :, :, g + 1:2 * g + 1], (3,))", "mass"),
    # the right wall mirrored like the left one
    ("f[:, :, n - g:] = torch.flip(f[:, :, n - 2 * g:n - g], (2,))",
     "f[:, :, n - g:] = f[:, :, n - 2 * g:n - g]", "symmetry"),
])
def test_mirror_pins_catch_planted_mistakes(monkeypatch, old, new, pin):
    from tests.test_oracle_pins import _mutant
    monkeypatch.setattr(bo.OracleSolver, "_mirror", _mutant(bo.OracleSolver._mirror, old, new))
    with pytest.raises(AssertionError):
        if pin == "mass":
            test_mirror_walls_conserve_mass("ppm")
        else:
            test_mirror_keeps_mirror_symmetric_state_symmetric()


def test_dirichlet_pin_catches_inflow_on_the_wrong_sign(monkeypatch):
    from tests.test_oracle_pins import _mutant
    monkeypatch.setattr(bo.OracleSolver, "_dirichlet", _mutant(
        bo.OracleSolver._dirichlet, "torch.where(C_q1 > 0,", "torch.where(C_q1 < 0,"))
    with pytest.raises(AssertionError):
        test_dirichlet_piecewise_constant_is_upwind_with_inflow_values()


def test_dirichlet_ghost_zone_is_inflow_values_and_zero_outflow():
    """After the refresh, a dirichlet ghost zone holds the user's values where
    the characteristic enters and zero where it leaves: the 'ghosted' DMDA
    axis is never written by globalToLocal (PETSc DM_BOUNDARY_GHOSTED), and
    the local vector starts at zero.  No reference test covers this; it is
    read from communicate.py:8-38 and nonlinear_solver.py:177-201, 306-313."""
    c = case_bc(left="dirichlet", right="dirichlet", N_q1=8, N_q2=1, p_dim=1, N_p1=6,
                inflow=(1.4, 0.2, 0.0))
    o = oracle_solver(c)
    o._communicate_f()
    g, n = o.N_ghost, o.f.shape[2]
    v = o.p1_center.reshape(-1)
    user = drifting_maxwellian(v, 0 * v, 0 * v, 1, 1.4, 0.2, 0.0)
    zero = torch.zeros_like(user)
    for k in range(g):
        assert torch.equal(o.f[:, 0, k, 0], torch.where(v > 0, user, zero))
        assert torch.equal(o.f[:, 0, n - 1 - k, 0], torch.where(v < 0, user, zero))


def test_ghosted_fill_pin_catches_periodic_ghosts(monkeypatch):
    """A ghosted axis filled periodically instead of left at zero."""
    monkeypatch.setattr(bo, "ghosted_fill", lambda x, g1, g2, p1, p2: bo.periodic_fill(x, g1, g2))
    with pytest.raises(AssertionError):
        test_dirichlet_ghost_zone_is_inflow_values_and_zero_outflow()
