"""The adjoint stage's per-cell code (bolt_amd/csrc/adjoint.h, run on the host
through tests/hostcheck) against torch autograd through the FP64 oracle's
df_dt_fvm: J(u)^T mu and sum mu dL/dtau for every reconstruction, both memory
layouts, with and without the BGK source, 1, 2 and 3 velocity dimensions.
States carry 20 % random noise: on smooth data the derivative of the WENO
weights inherits the rounding-dominated expanded indicator (the forward
step's conditioning), which would make any comparison a test of rounding."""
import ctypes
import math

import numpy as np
import pytest
import torch
import torch.nn.functional as Fn

from oracle import bolt_oracle as bo
from tests.cases import maxwellian_case, oracle_solver
from tests.hostcheck import harness as H

F64 = torch.float64


def _host():
    lib = H.lib()
    lib.host_adjoint_stage.argtypes = [ctypes.POINTER(H.L.Grid), ctypes.POINTER(H.L.Physics),
                                       ctypes.POINTER(H.L.Adjoint)]
    lib.host_adjoint_stage.restype = ctypes.c_int
    return lib


def _moment_block(o, u_int):
    """[12][Q] moment block of an interior array in the kernels' convention:
    raw moments (x dV), then n (m/2 pi k T)^(d/2), u1, u2, u3, T, n, m/2kT
    from the oracle's own BGK_moments (collision_operator.py:102-116)."""
    m5 = H.moments_np(o, u_int)
    n, m1, m2, m3, E = (torch.from_numpy(m5[k]) for k in range(5))
    p = o.params
    mass, kB, d = float(o.mass.reshape(-1)[0]), float(p.boltzmann_constant), p.p_dim
    ns = types_ns(mass)
    u1, u2, u3, T = bo.BGK_moments(n, m1, m2, m3, E, ns(p.p_dim))
    pref = n * (mass / (2 * math.pi * kB * T)) ** (d / 2)
    rows = [n, m1, m2, m3, E, pref, u1, u2, u3, T, n, mass / (2 * kB * T)]
    return np.ascontiguousarray(torch.stack(rows).numpy())


def types_ns(mass):
    import types

    def make(p_dim):
        return types.SimpleNamespace(mass=mass, p_dim=p_dim)
    return make


def _torch_vjp(o, u_int, mu_int, tau_value):
    """d/du and d/dtau of sum(mu * L(u)) with autograd through the oracle."""
    tau = torch.tensor(tau_value, dtype=F64, requires_grad=True)
    o.params.tau = lambda q1, q2, p1, p2, p3: tau * (q1 * 0 + 1) * (p1 * 0 + 1)
    u = u_int.clone().requires_grad_(True)
    g1, g2 = o.N_g1, o.N_g2
    f = bo.periodic_fill(Fn.pad(u, (g2, g2, g1, g1)), g1, g2)
    L = o.interior(o.df_dt_fvm(f))
    gu, gt = torch.autograd.grad((L * mu_int).sum(), [u, tau], allow_unused=True)
    return gu, (0.0 if gt is None else float(gt))


CASES = {
    "2d2v": dict(N_q1=7, N_q2=6, N_p1=6, N_p2=5, p_dim=2),
    "1d1v": dict(N_q1=9, N_q2=1, N_p1=10, p_dim=1),
    "2d3v": dict(N_q1=5, N_q2=6, N_p1=4, N_p2=4, N_p3=3, p_dim=3, p_max=6.0),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("rq", ["piecewise-constant", "minmod", "ppm", "weno5"])
@pytest.mark.parametrize("source", [True, False])
@pytest.mark.parametrize("layout", [0, 1])
def test_adjoint_stage_matches_autograd(name, rq, source, layout):
    kw = dict(CASES[name], recon_q=rq, recon_p=rq, riemann="lax-friedrichs", tau=0.03,
              source=source, modes=((1, 1),), amps=(3e-2,), phases=(0.4,))
    c = maxwellian_case(**kw)
    o = oracle_solver(c)
    gen = torch.Generator().manual_seed(7)
    u = o.interior().clone() * (1 + 0.2 * torch.rand(o.interior().shape, generator=gen,
                                                      dtype=F64))
    mu = torch.randn(u.shape, generator=gen, dtype=F64)
    gu, gt = _torch_vjp(o, u, mu, 0.03)
    g, ph = H.grid_physics(o)
    g.layout = layout
    perm = (2, 1, 0) if layout == 0 else (2, 0, 1)   # (P, 1, N1, N2) -> storage
    store = lambda x: np.ascontiguousarray(x[:, 0].permute(*perm).numpy())
    u_s, mu_s = store(u), store(mu)
    mom = _moment_block(o, np.ascontiguousarray(u[:, 0].permute(2, 1, 0).numpy()))
    out = np.empty_like(u_s)
    Q = o.N_q1 * o.N_q2
    ws = np.zeros(5 * Q)
    tq = np.zeros(Q)
    P = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    adj = H.L.Adjoint(P(u_s), P(mom), P(mu_s), None, P(out), 0.0, 0.0, 1.0, P(tq), P(ws))
    _host().host_adjoint_stage(ctypes.byref(g), ctypes.byref(ph), ctypes.byref(adj))
    want = store(gu)
    err = np.abs(out - want).max() / np.abs(want).max()
    assert err <= 1e-10, err
    if source:
        assert abs(tq.sum() - gt) <= 1e-10 * abs(gt), (tq.sum(), gt)


def test_adjoint_stage_combination():
    """out = a*base + b*mu + c*J^T mu."""
    c = maxwellian_case(**dict(CASES["2d2v"], recon_q="weno5", recon_p="weno5", tau=0.05))
    o = oracle_solver(c)
    gen = torch.Generator().manual_seed(3)
    u = o.interior().clone() * (1 + 0.2 * torch.rand(o.interior().shape, generator=gen,
                                                      dtype=F64))
    mu = torch.randn(u.shape, generator=gen, dtype=F64)
    base = torch.randn(u.shape, generator=gen, dtype=F64)
    gu, _ = _torch_vjp(o, u, mu, 0.05)
    g, ph = H.grid_physics(o)
    store = lambda x: np.ascontiguousarray(x[:, 0].permute(2, 1, 0).numpy())
    u_s, mu_s, b_s = store(u), store(mu), store(base)
    mom = _moment_block(o, u_s)
    out = np.empty_like(u_s)
    ws = np.zeros(5 * o.N_q1 * o.N_q2)
    P = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    adj = H.L.Adjoint(P(u_s), P(mom), P(mu_s), P(b_s), P(out), 0.3, 0.7, 0.01, None, P(ws))
    _host().host_adjoint_stage(ctypes.byref(g), ctypes.byref(ph), ctypes.byref(adj))
    want = 0.3 * b_s + 0.7 * mu_s + 0.01 * store(gu)
    assert np.abs(out - want).max() <= 1e-11 * np.abs(want).max()
