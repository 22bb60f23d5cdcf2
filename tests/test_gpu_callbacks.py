"""GPU: user plugins through the solver (C_q, C_p, source and a varying tau
evaluated with torch on the device each RHS and fed to the generic stage
kernel through boltx_fvm_stage_coeffs), per step against the oracle with the
same callbacks; the reference's own plugin modules (matched by name) keep the
fused kernels."""
import numpy as np
import pytest
import torch

from bolt_amd import _lib
from tests.cases import hip_solver, oracle_solver
from tests.test_callbacks_cpu import _case

pytestmark = pytest.mark.gpu


def _per_step(c, steps, tol):
    o, nls = oracle_solver(c), hip_solver(c)
    assert nls._general and nls._layout == _lib.LAYOUT["reference"]
    worst = 0.0
    for _ in range(steps):
        nls.f = o.f.to(nls.device)
        nls.time_elapsed = o.time_elapsed
        o.strang_timestep(c.dt)
        nls.strang_timestep(c.dt)
        g1, g2 = nls.N_g1, nls.N_g2
        a = nls.f[:, :, g1:g1 + nls.N_q1, g2:g2 + nls.N_q2].cpu().numpy()
        b = o.interior().numpy()
        worst = max(worst, np.abs(a - b).max() / np.abs(b).max())
    print(f"per-step rel Linf worst {worst:.3e}", flush=True)
    assert worst <= tol


@pytest.mark.parametrize("kind", ["cq", "cp", "source", "tau"])
@pytest.mark.parametrize("recon,riemann", [("minmod", "upwind-flux"), ("weno5", "lax-friedrichs")])
@pytest.mark.parametrize("stepper", ["RK2", "SSP-RK3"])
def test_user_terms_per_step(gpu, kind, recon, riemann, stepper):
    c = _case(kind, recon, riemann)
    c.params.time_stepper_fvm = stepper
    c.dt = 1e-3
    _per_step(c, 3, 1e-12)


def test_user_terms_fp32_bulk(gpu):
    """FP32-bulk with every user term at once (C_q, C_p with electrostatic
    fields, source): rel-L2 against the FP64 oracle after 5 steps."""
    c64 = _case("cp", "weno5", "lax-friedrichs")
    c32 = _case("cp", "weno5", "lax-friedrichs")
    from tests.test_callbacks_cpu import user_source
    for c in (c64, c32):
        c.source = user_source
    c32.params.precision = "fp32-bulk"
    o, nls = oracle_solver(c64), hip_solver(c32)
    for _ in range(5):
        o.strang_timestep(1e-3)
        nls.strang_timestep(1e-3)
    g1, g2 = nls.N_g1, nls.N_g2
    a = nls.f[:, :, g1:g1 + nls.N_q1, g2:g2 + nls.N_q2].double().cpu()
    b = o.interior()
    rel = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
    print(f"user terms FP32-bulk 5 steps: rel L2 {rel:.3e}", flush=True)
    assert rel <= 1e-5


def test_reference_plugin_modules_keep_the_fused_path(gpu):
    """Modules named as the reference's plugins (bolt.src.
    nonrelativistic_boltzmann.*) holding the same functions are recognised
    by name: the solver keeps the fused q1-line kernels and matches the run
    with this package's own plugin modules bit for bit."""
    import types
    import bolt_amd.src.nonrelativistic_boltzmann.advection_terms as adv
    import bolt_amd.src.nonrelativistic_boltzmann.collision_operator as coll
    from bolt_amd.physical_system import physical_system
    from bolt_amd.nonlinear_solver import nonlinear_solver
    from tests.cases import case_2d2v

    def renamed(fn, module):
        g = types.FunctionType(fn.__code__, fn.__globals__, fn.__name__, fn.__defaults__,
                               fn.__closure__)
        g.__module__ = module
        return g
    ref_adv = types.ModuleType("bolt.src.nonrelativistic_boltzmann.advection_terms")
    for n in ("A_q", "C_q", "A_p", "C_p"):
        setattr(ref_adv, n, renamed(getattr(adv, n), ref_adv.__name__))
    bgk = renamed(coll.BGK, "bolt.src.nonrelativistic_boltzmann.collision_operator")
    c = case_2d2v(seed=1, N_q=64, N_p=8)
    a = hip_solver(c)
    import bolt_amd.src.nonrelativistic_boltzmann.moments as mom
    b = nonlinear_solver(physical_system(c.domain, c.boundary_conditions, c.params,
                                         c.initialize, ref_adv, bgk, mom))
    assert not b._general and b._layout == _lib.LAYOUT["q1-lines"]
    for _ in range(2):
        a.strang_timestep(c.dt)
        b.strang_timestep(c.dt)
    assert torch.equal(a._f, b._f)
