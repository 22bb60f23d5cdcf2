"""The adjoint on the GPU (through the C-ABI): stage vector-Jacobian
products against autograd through the FP64 oracle, the checkpointed
whole-run gradient against autograd through plain-torch oracle steps, and
the adjoint workload's gradients against central finite differences."""
import ctypes
import math

import numpy as np
import pytest
import torch
import torch.nn.functional as Fn

from oracle import bolt_oracle as bo
from tests.cases import case_c4, hip_solver, maxwellian_case, maxwellian_interior, oracle_solver

pytestmark = pytest.mark.gpu
F64 = torch.float64


def _oracle_vjp(o, u_int, mu_int, tau_value):
    tau = torch.tensor(tau_value, dtype=F64, requires_grad=True)
    o.params.tau = lambda q1, q2, p1, p2, p3: tau * (q1 * 0 + 1) * (p1 * 0 + 1)
    u = u_int.clone().requires_grad_(True)
    g1, g2 = o.N_g1, o.N_g2
    f = bo.periodic_fill(Fn.pad(u, (g2, g2, g1, g1)), g1, g2)
    L = o.interior(o.df_dt_fvm(f))
    gu, gt = torch.autograd.grad((L * mu_int).sum(), [u, tau])
    return gu[:, 0], float(gt)


def _ws_bytes(s):
    return int(s._L.boltx_adjoint_workspace_size(ctypes.byref(s._grid), ctypes.byref(s._phys)))


@pytest.mark.parametrize("layout", ["reference", "q1-lines"])
@pytest.mark.parametrize("rq", ["minmod", "ppm", "weno5"])
def test_stage_adjoint_vs_oracle_autograd(gpu, layout, rq):
    from bolt_amd import adjoint as A
    c = maxwellian_case(N_q1=64 if layout == "q1-lines" else 12, N_q2=6, N_p1=8, N_p2=6,
                        p_dim=2, recon_q=rq, recon_p=rq, riemann="lax-friedrichs", tau=0.04,
                        memory_layout=layout, modes=((1, 1),), amps=(3e-2,), phases=(0.3,))
    o = oracle_solver(c)
    gen = torch.Generator().manual_seed(11)
    u = o.interior().clone() * (1 + 0.2 * torch.rand(o.interior().shape, generator=gen, dtype=F64))
    mu = torch.randn(u.shape, generator=gen, dtype=F64)
    gu, gt = _oracle_vjp(o, u, mu, 0.04)
    s = hip_solver(c)
    U = A.interior_to_internal(s, u[:, 0].to(gpu))
    MU = A.interior_to_internal(s, mu[:, 0].to(gpu))
    mom = A._moments_of(s, U)
    ws = torch.empty(_ws_bytes(s), dtype=torch.uint8, device=gpu)
    tq = torch.empty(s._Q, dtype=F64, device=gpu)
    out = A._adjoint_stage(s, U, mom, MU, None, 0.0, 0.0, 1.0, ws, tq)
    got = A.internal_to_interior(s, out).cpu()
    err = float((got - gu).abs().max() / gu.abs().max())
    terr = abs(float(tq.sum()) - gt) / abs(gt)
    print(f"{layout} {rq}: J^T mu rel Linf {err:.3e}, dL/dtau rel {terr:.3e}", flush=True)
    assert err <= 1e-10 and terr <= 1e-10


@pytest.mark.parametrize("n1,n2", [(128, 5), (64, 37)])
def test_line_adjoint_shapes_and_second_base(gpu, n1, n2):
    """The reverse-mode line kernel (q1-line layout, WENO5) on two q1 cells
    per lane and on a q2 extent that splits into uneven segments, against
    oracle autograd; the second base term adds a2 * base2."""
    from bolt_amd import adjoint as A
    c = maxwellian_case(N_q1=n1, N_q2=n2, N_p1=6, N_p2=4, p_dim=2, recon_q="weno5",
                        recon_p="weno5", riemann="lax-friedrichs", tau=0.04,
                        memory_layout="q1-lines", modes=((1, 1),), amps=(3e-2,), phases=(0.7,))
    o = oracle_solver(c)
    gen = torch.Generator().manual_seed(5)
    u = o.interior().clone() * (1 + 0.2 * torch.rand(o.interior().shape, generator=gen, dtype=F64))
    mu = torch.randn(u.shape, generator=gen, dtype=F64)
    gu, gt = _oracle_vjp(o, u, mu, 0.04)
    s = hip_solver(c)
    U = A.interior_to_internal(s, u[:, 0].to(gpu))
    MU = A.interior_to_internal(s, mu[:, 0].to(gpu))
    mom = A._moments_of(s, U)
    ws = torch.empty(_ws_bytes(s), dtype=torch.uint8, device=gpu)
    tq = torch.empty(s._Q, dtype=F64, device=gpu)
    out = A._adjoint_stage(s, U, mom, MU, None, 0.0, 0.0, 1.0, ws, tq)
    got = A.internal_to_interior(s, out).cpu()
    err = float((got - gu).abs().max() / gu.abs().max())
    terr = abs(float(tq.sum()) - gt) / abs(gt)
    print(f"n1={n1} n2={n2}: J^T mu rel Linf {err:.3e}, dL/dtau rel {terr:.3e}", flush=True)
    assert err <= 1e-10 and terr <= 1e-10
    B1 = torch.randn(U.shape, generator=gen, dtype=F64).to(gpu)
    B2 = torch.randn(U.shape, generator=gen, dtype=F64).to(gpu)
    one = A._adjoint_stage(s, U, mom, MU, B1, 0.3, 0.7, 0.01, ws, None)
    two = A._adjoint_stage(s, U, mom, MU, B1, 0.3, 0.7, 0.01, ws, None, base2=B2, a2=-1.25)
    d = float((two - one - (-1.25) * B2).abs().max() / two.abs().max())
    assert d <= 1e-14, d


@pytest.mark.parametrize("layout,n1,np1,np2", [("q1-lines", 40, 12, 10),
                                               ("reference", 10, 20, 15)])
def test_stage_adjoint_ragged_reduction(gpu, layout, n1, np1, np2):
    """The p-parallel BGK reduction of the adjoint with q1 lines that leave
    lanes of a 64-wide block empty (n1 = 40) and with more p-cells than the
    256 threads of a q-cell's workgroup (P = 300)."""
    from bolt_amd import adjoint as A
    c = maxwellian_case(N_q1=n1, N_q2=7, N_p1=np1, N_p2=np2, p_dim=2, recon_q="weno5",
                        recon_p="weno5", riemann="lax-friedrichs", tau=0.04, memory_layout=layout,
                        modes=((1, 1),), amps=(3e-2,), phases=(0.3,))
    o = oracle_solver(c)
    gen = torch.Generator().manual_seed(12)
    u = o.interior().clone() * (1 + 0.2 * torch.rand(o.interior().shape, generator=gen, dtype=F64))
    mu = torch.randn(u.shape, generator=gen, dtype=F64)
    gu, gt = _oracle_vjp(o, u, mu, 0.04)
    s = hip_solver(c)
    U = A.interior_to_internal(s, u[:, 0].to(gpu))
    MU = A.interior_to_internal(s, mu[:, 0].to(gpu))
    mom = A._moments_of(s, U)
    ws = torch.empty(_ws_bytes(s), dtype=torch.uint8, device=gpu)
    tq = torch.empty(s._Q, dtype=F64, device=gpu)
    out = A._adjoint_stage(s, U, mom, MU, None, 0.0, 0.0, 1.0, ws, tq)
    got = A.internal_to_interior(s, out).cpu()
    err = float((got - gu).abs().max() / gu.abs().max())
    terr = abs(float(tq.sum()) - gt) / abs(gt)
    print(f"{layout} n1={n1} P={np1 * np2}: J^T mu rel Linf {err:.3e}, dL/dtau rel {terr:.3e}",
          flush=True)
    assert err <= 1e-10 and terr <= 1e-10
    out2 = A._adjoint_stage(s, U, mom, MU, None, 0.0, 0.0, 1.0, ws, tq)
    assert torch.equal(out, out2)  # fixed summation order: bit-reproducible


def _c4_loss(n, E, dq):
    return ((n - 1) ** 2).sum() * dq + E.sum() * dq


@pytest.mark.parametrize("stepper", ["SSP-RK3", "RK2"])
def test_evolve_gradient_vs_oracle_autograd(gpu, stepper):
    """Whole-run gradients (checkpointed adjoint) against autograd through the
    plain-torch oracle steps: d loss / d dn and d loss / d tau."""
    from bolt_amd import adjoint as A
    c = case_c4(seed=1, N_q=12, N_p=10, stepper=stepper)
    steps, every, tau0 = 8, 3, 0.05
    # oracle: autograd through SSP-RK3 steps
    o = oracle_solver(c)
    dn = c.dn.clone().requires_grad_(True)
    tau = torch.tensor(tau0, dtype=F64, requires_grad=True)
    o.params.tau = lambda q1, q2, p1, p2, p3: tau * (q1 * 0 + 1) * (p1 * 0 + 1)
    v = [x.reshape(-1) for x in (o.p1_center, o.p2_center, o.p3_center)]
    f0 = maxwellian_interior(*v, 2, dn)
    g1, g2 = o.N_g1, o.N_g2
    o.f = bo.periodic_fill(Fn.pad(f0.unsqueeze(1), (g2, g2, g1, g1)), g1, g2)
    for _ in range(steps):
        if stepper == "RK2":
            o.timestep_fvm_rk2(c.dt)
        else:
            o.timestep_fvm_ssprk3(c.dt)
    fT = o.interior()[:, 0]
    dv = float((o.dp1 * o.dp2 * o.dp3).reshape(-1)[0])
    n = fT.sum(0) * dv
    E = (0.5 * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2)).reshape(-1, 1, 1).mul(fT).sum(0) * dv
    J = _c4_loss(n, E, o.dq1 * o.dq2)
    g_dn_ref, g_tau_ref = torch.autograd.grad(J, [dn, tau])
    # HIP: evolve + the adjoint
    s = hip_solver(c)
    dn_h = c.dn.clone().to(gpu).requires_grad_(True)
    tau_h = torch.tensor(tau0, dtype=F64, device=gpu, requires_grad=True)
    vh = [x.reshape(-1) for x in (s.p1_center, s.p2_center, s.p3_center)]
    fT_h = A.evolve(s, maxwellian_interior(*vh, 2, dn_h), tau_h, steps, c.dt, every)
    J_h = _c4_loss(A.moment(s, fT_h, "density"), A.moment(s, fT_h, "energy"), s.dq1 * s.dq2)
    g_dn, g_tau = torch.autograd.grad(J_h, [dn_h, tau_h])
    e_j = abs(float(J_h) - float(J)) / abs(float(J))
    e_dn = float((g_dn.cpu() - g_dn_ref).abs().max() / g_dn_ref.abs().max())
    e_tau = abs(float(g_tau) - float(g_tau_ref)) / abs(float(g_tau_ref))
    print(f"loss rel {e_j:.3e}; d/d dn rel Linf {e_dn:.3e}; d/d tau rel {e_tau:.3e}", flush=True)
    assert e_j <= 1e-12 and e_dn <= 1e-8 and e_tau <= 1e-8


def test_c4_gradients_vs_central_differences(gpu):
    """The adjoint workload at full size (64^2 x 32^2, 100 SSP-RK3 steps,
    checkpoints every 10): directional derivatives of the loss along 32
    random unit directions in dn, and its derivative in tau, against central
    differences with h = 1e-4 in FP64; agreement to 1e-5 relative.

    The loss is J1 + J2 with J1 = int (n - 1)^2 dq and J2 the total kinetic
    energy.  The scheme conserves J2 exactly for advection (periodic fluxes
    telescope) and to quadrature error (~e^-50 here) for BGK, so J2 does not
    depend on tau: d/d tau is checked on J1, whose difference quotient is not
    buried under the rounding of J2 ~ 1, and the adjoint's d J2 / d tau must
    vanish to rounding."""
    from bolt_amd import adjoint as A
    c = case_c4(seed=0)
    s = hip_solver(c)
    vh = [x.reshape(-1) for x in (s.p1_center, s.p2_center, s.p3_center)]
    dq = s.dq1 * s.dq2

    def parts(dn, tau):
        fT = A.evolve(s, maxwellian_interior(*vh, 2, dn), tau, c.steps, c.dt, c.checkpoint_every)
        n, E = A.moment(s, fT, "density"), A.moment(s, fT, "energy")
        return ((n - 1) ** 2).sum() * dq, E.sum() * dq

    dn = c.dn.clone().to(gpu).requires_grad_(True)
    tau = torch.tensor(0.05, dtype=F64, device=gpu, requires_grad=True)
    J1, J2 = parts(dn, tau)
    g_dn, g_tau = torch.autograd.grad(J1 + J2, [dn, tau], retain_graph=True)
    (g1_tau,) = torch.autograd.grad(J1, [tau], retain_graph=True)
    (g2_tau,) = torch.autograd.grad(J2, [tau])
    h = 1e-4
    gen = torch.Generator().manual_seed(5)
    worst = 0.0
    with torch.no_grad():
        base = c.dn.to(gpu)
        for k in range(32):
            d = torch.randn(base.shape, generator=gen, dtype=F64).to(gpu)
            d = d / torch.linalg.vector_norm(d)
            jp, jm = parts(base + h * d, tau), parts(base - h * d, tau)
            fd = (float(jp[0] + jp[1]) - float(jm[0] + jm[1])) / (2 * h)
            ad = float((g_dn * d).sum())
            worst = max(worst, abs(fd - ad) / abs(ad))
        fd1 = (float(parts(base, tau + h)[0]) - float(parts(base, tau - h)[0])) / (2 * h)
    e_tau = abs(fd1 - float(g1_tau)) / abs(float(g1_tau))
    print(f"C4: worst directional rel err {worst:.3e}; d J1/d tau rel err {e_tau:.3e} "
          f"(AD {float(g1_tau):.6e}, FD {fd1:.6e}); d J2/d tau {float(g2_tau):.3e}; "
          f"d J/d tau {float(g_tau):.6e}", flush=True)
    assert worst <= 1e-5 and e_tau <= 1e-5
    assert abs(float(g2_tau)) <= 1e-4 * abs(float(g1_tau))
