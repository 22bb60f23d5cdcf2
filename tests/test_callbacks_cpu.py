"""User plugins (C_q, C_p, source, a varying tau) through the generic stage
code with caller-supplied coefficient arrays, on the CPU harness: one RHS
evaluation u + dt L(u) against the oracle's df_dt_fvm with the same
callbacks (df_dt_fvm.py:67-97, 251-323; collision_operator.py:119-128).

The callbacks are written with torch ops, as a user of this package writes
them, so the same functions drive the oracle (CPU tensors) and the solver
(device tensors).  Each case gives its terms a dependence the built-in ones
lack (q-dependent C_q, p1-dependent C_p1, a source reading moments, tau over
q and p), so a face velocity read at the wrong face or cell fails the bar."""
import math
import types

import numpy as np
import pytest
import torch

from tests.cases import maxwellian_case, oracle_solver
from tests.hostcheck import harness as H

L = H.L  # the harness's ctypes structs

TWO_PI = 2 * math.pi


def user_advection():
    """C_q varying over q (at the face positions the solver passes), C_p
    with a q-dependent force and a drag along each p axis (so C_p1 differs
    between the two p1-faces of a cell)."""
    def C_q(t, q1, q2, v1, v2, v3, params):
        return (v1 * (1 + 0.3 * torch.sin(TWO_PI * q1)) + 0 * q2,
                v2 * (0.8 + 0.2 * torch.cos(TWO_PI * q2)) + 0.1 * torch.sin(TWO_PI * q1))

    def C_p(t, q1, q2, v1, v2, v3, fields_solver, params):
        E1, E2, E3, B1, B2, B3 = fields_solver.get_fields()
        qm = params.charge / params.mass
        return (qm * E1 - 0.05 * v1 + 0.2 * torch.sin(TWO_PI * q1),
                qm * E2 - 0.05 * v2 + 0.1 * torch.cos(TWO_PI * q2),
                0 * v3 + 0 * q1)
    m = types.ModuleType("user_advection")
    m.A_q, m.C_q, m.A_p, m.C_p = C_q, C_q, C_p, C_p
    return m


def user_source(f, t, q1, q2, v1, v2, v3, moments, params, flag=False):
    """A q-dependent damping plus a term fed by the density moment."""
    n = moments("density", f)
    return -0.5 * f * torch.cos(TWO_PI * q1) + 1e-3 * n * torch.exp(-(v1**2 + v2**2) / 2)


def varying_tau(q1, q2, p1, p2, p3):
    return 0.01 * (1 + 0.5 * torch.sin(TWO_PI * q1)) * (1 + 0.01 * p1**2) + 0 * q2 + 0 * p2


def _case(kind, recon, riemann):
    kw = dict(N_q1=8, N_q2=6, N_p1=8, N_p2=6, N_ghost=3, p_max=6.0, p_dim=2,
              modes=((1, 1), (2, 1)), amps=(3e-2, 2e-2), phases=(0.3, 1.1), recon_q=recon,
              recon_p=recon, riemann=riemann, stepper="SSP-RK3", tau=0.02)
    if kind == "cp":
        kw.update(fields=True, fields_type="electrostatic", fields_initialize="fft")
    c = maxwellian_case(**kw)
    if kind in ("cq", "cp"):
        c.advection = user_advection()
    if kind == "source":
        c.source = user_source
    if kind == "tau":
        c.params.tau = varying_tau
    return c


def _compare(kind, recon, riemann, at_faces=True):
    c = _case(kind, recon, riemann)
    o = oracle_solver(c)
    dt = 1e-3
    o.dt = dt
    o._communicate_f()
    f = o.f
    want = H.to_internal(o, f + o.df_dt_fvm(f) * dt)
    keep = []
    co = L.Coeffs()
    t, p = o.time_elapsed, o.params
    adv = getattr(c, "advection", None)
    padv = None
    q1f = o.q1_left_center if at_faces else o.q1_center
    p1f = o.p1_left if at_faces else o.p1_center
    if kind in ("cq", "cp"):
        co.c_q1 = H.coeff(o, adv.C_q(t, q1f, o.q2_left_center, o.p1_center,
                                     o.p2_center, o.p3_center, p)[0], keep)
        co.c_q2 = H.coeff(o, adv.C_q(t, o.q1_center_bot, o.q2_center_bot, o.p1_center,
                                     o.p2_center, o.p3_center, p)[1], keep)
    if kind == "cp":
        # the electrostatic fields of f are those df_dt_fvm just solved for
        fs = o.fields_solver
        co.c_p1 = H.coeff(o, adv.C_p(t, o.q1_center, o.q2_center, p1f, o.p2_left,
                                     o.p3_left, fs, p)[0], keep)
        co.c_p2 = H.coeff(o, adv.C_p(t, o.q1_center, o.q2_center, o.p1_bottom, o.p2_bottom,
                                     o.p3_bottom, fs, p)[1], keep)
        co.c_p3 = H.coeff(o, adv.C_p(t, o.q1_center, o.q2_center, o.p1_back, o.p2_back,
                                     o.p3_back, fs, p)[2], keep)
        padv = 1
    if kind == "source":
        co.source = H.coeff(o, c.source(f, t, o.q1_center, o.q2_center, o.p1_center,
                                        o.p2_center, o.p3_center, o.compute_moments, p, False),
                            keep)
    if kind == "tau":
        co.tau = H.coeff(o, varying_tau(o.q1_center, o.q2_center, o.p1_center, o.p2_center,
                                        o.p3_center), keep)
    u = H.to_internal(o, f)
    got, _ = H.stage_coeffs(o, u, u, None, 0.0, 1.0, dt, co, padv=padv)
    return np.abs(got - want).max() / np.abs(want).max()


@pytest.mark.parametrize("kind", ["cq", "cp", "source", "tau"])
@pytest.mark.parametrize("recon,riemann", [("minmod", "upwind-flux"), ("weno5", "lax-friedrichs"),
                                           ("ppm", "upwind-flux")])
def test_user_terms_match_oracle(kind, recon, riemann):
    err = _compare(kind, recon, riemann)
    assert err <= 1e-13, err


def test_user_terms_catch_misplaced_faces(monkeypatch):
    """The bar sees a face velocity read one cell off (every coefficient
    read one q1-cell to the right) and callbacks evaluated at cell centres
    instead of the faces (C_q1 at q1_center, C_p1 at p1_center)."""
    orig = H.coeff

    def shifted(o, x, keep):
        c = orig(o, x, keep)
        c.ptr += 8 * c.stride_q1
        return c
    monkeypatch.setattr(H, "coeff", shifted)
    assert _compare("cq", "minmod", "upwind-flux") > 1e-6
    monkeypatch.undo()
    assert _compare("cq", "minmod", "upwind-flux", at_faces=False) > 1e-6
    assert _compare("cp", "minmod", "upwind-flux", at_faces=False) > 1e-6


def test_plugin_identity_by_name():
    """The built-in plugins are recognised as themselves or by the dotted
    name of their module (the reference's bolt.src.nonrelativistic_boltzmann
    modules); a user function of another module, or of another name, is a
    user callback."""
    from bolt_amd.nonlinear_solver import _is_builtin, _is_builtin_module
    import bolt_amd.src.nonrelativistic_boltzmann.advection_terms as adv
    import bolt_amd.src.nonrelativistic_boltzmann.moments as mom

    def C_q(t, q1, q2, v1, v2, v3, params):
        return (v1, v2)
    assert _is_builtin(adv.C_q, adv.C_q, "advection_terms")
    C_q.__module__ = "bolt.src.nonrelativistic_boltzmann.advection_terms"
    assert _is_builtin(C_q, adv.C_q, "advection_terms")
    C_q.__module__ = "nonrelativistic_boltzmann.advection_terms"
    assert _is_builtin(C_q, adv.C_q, "advection_terms")
    C_q.__module__ = "my_physics.advection_terms"
    assert not _is_builtin(C_q, adv.C_q, "advection_terms")
    C_q.__module__ = "bolt.src.nonrelativistic_boltzmann.advection_terms"
    C_q.__name__ = "C_q_relativistic"
    assert not _is_builtin(C_q, adv.C_q, "advection_terms")
    assert not _is_builtin(user_advection().C_q, adv.C_q, "advection_terms")
    m = types.ModuleType("bolt.src.nonrelativistic_boltzmann.moments")
    assert _is_builtin_module(m, mom) and _is_builtin_module(mom, mom)
    assert not _is_builtin_module(types.ModuleType("my_moments"), mom)
