"""GPU: the electrodynamic coupling of the HIP path (through the C-ABI).

- the k = 0 Langmuir oscillation of a uniform plasma follows the stepper's own
  scalar recurrence (the reference's field coupling for RK2, the once-per-step
  coupling for SSP-RK3) and gyration / acceleration in uniform fields follow
  the exact discrete rotation / drift, independently of the oracle;
- per-step parity of the 2D3V electromagnetic step against the oracle for both
  couplings, including the field time levels the SSP-RK3 stages see.
"""
import math

import numpy as np
import pytest
import torch

from tests.cases import case_2d3v, hip_solver, oracle_solver
from tests.test_oracle_fields import langmuir_recurrence, uniform_plasma_case

pytestmark = pytest.mark.gpu


def _bulk(nls):
    """(n, n u1, n u2, E1) at q-cell 0 (the state is q-uniform)."""
    g1, g2 = nls.N_g1, nls.N_g2
    n = float(nls.compute_moments("density")[0, 0, g1, g2])
    m1 = float(nls.compute_moments("mom_v1_bulk")[0, 0, g1, g2])
    m2 = float(nls.compute_moments("mom_v2_bulk")[0, 0, g1, g2])
    E1 = float(nls.fields_solver._yee[0, 0])
    return n, m1, m2, E1


@pytest.mark.parametrize("precision", ["fp64", "fp32-bulk"])
@pytest.mark.parametrize("stepper", ["RK2", "SSP-RK3"])
def test_langmuir_k0_recurrence(gpu, stepper, precision):
    dt, n_steps = 0.05, 40
    c = uniform_plasma_case(stepper=stepper, u0=(0.05,), eps=1.0, N_p=48, p_dim=1)
    c.params.precision = precision
    nls = hip_solver(c)
    traj = [_bulk(nls)]
    for _ in range(n_steps):
        nls.strang_timestep(dt)
        traj.append(_bulk(nls))
    n, m0, _, E0 = traj[0]
    rec = langmuir_recurrence(stepper, dt, n_steps, n, m0, E0)
    sm = max(abs(t[1]) for t in traj)
    sE = max(abs(t[3]) for t in traj)
    tol = 1e-12 if precision == "fp64" else 2e-6
    worst = max(max(abs(t[1] - r[0]) / sm, abs(t[3] - r[1]) / sE) for t, r in zip(traj, rec))
    print(f"Langmuir k=0 {stepper} {precision}: worst deviation from the recurrence "
          f"{worst:.3e}", flush=True)
    assert worst <= tol


@pytest.mark.parametrize("stepper", ["RK2", "SSP-RK3"])
def test_gyration_exact_discrete_rotation(gpu, stepper):
    qm, B0, u0, dt, n_steps = -1.0, 0.5, 0.6, 0.02, 100
    c = uniform_plasma_case(stepper=stepper, u0=(u0, 0.0), B0=(0.0, 0.0, B0), eps=1e12,
                            N_p=64, p_max=16.0)
    nls = hip_solver(c)
    for _ in range(n_steps):
        nls.strang_timestep(dt)
    n, m1, m2, _ = _bulk(nls)
    z = -1j * qm * B0 * dt
    R = 1 + z + z * z / 2 + (z ** 3 / 6 if stepper == "SSP-RK3" else 0)
    want = u0 * R ** n_steps
    err = abs(complex(m1 / n, m2 / n) - want)
    print(f"gyration {stepper}: |u - u0 R^N| = {err:.3e}", flush=True)
    assert err <= 1e-11 * u0


@pytest.mark.parametrize("stepper", ["RK2", "SSP-RK3"])
def test_2d3v_per_step_with_fields(gpu, stepper):
    """Per step from the oracle's state (f reset each step, the fields carried
    by both): f to 1e-12, and for SSP-RK3 the cell-centred fields and both
    stage time levels (fields at n and at n+1/2) after every step."""
    c = case_2d3v(seed=9, N_q=8, N_p=6, stepper=stepper)
    o, nls = oracle_solver(c), hip_solver(c)
    g1, g2 = nls.N_g1, nls.N_g2
    worst_f = worst_fields = 0.0
    for _ in range(4):
        nls.f = o.f.to(nls.device)
        nls.time_elapsed = o.time_elapsed
        o.strang_timestep(c.dt)
        nls.strang_timestep(c.dt)
        a = nls.f[:, :, g1:g1 + nls.N_q1, g2:g2 + nls.N_q2].cpu().numpy()
        b = o.interior().numpy()
        worst_f = max(worst_f, float(np.abs(a - b).max() / np.abs(b).max()))
        fs, ofs = nls.fields_solver, o.fields_solver
        pairs = [(fs._cc, ofs.cell_centered_EM_fields)]
        if stepper == "SSP-RK3":
            pairs += [(fs._cc_at_n, ofs.cell_centered_EM_fields_at_n),
                      (fs._cc_at_n_plus_half, ofs.cell_centered_EM_fields_at_n_plus_half)]
        for mine, ref in pairs:
            r = ref[:, 0, g1:g1 + nls.N_q1, g2:g2 + nls.N_q2].permute(0, 2, 1).reshape(6, -1)
            scale = max(float(r.abs().max()), 1e-30)
            worst_fields = max(worst_fields,
                               float((mine.cpu() - r).abs().max()) / scale)
    print(f"2D3V {stepper} per step: f {worst_f:.3e}, fields {worst_fields:.3e}", flush=True)
    assert worst_f <= 1e-12
    assert worst_fields <= 1e-12
