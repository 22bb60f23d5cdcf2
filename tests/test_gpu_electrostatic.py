"""Electrostatic coupling (FFT Poisson each RHS, df_dt_fvm.py:235-242,
fft.py:9-78) on the GPU against the oracle: fields at initialisation and
per-step states from the identical input."""
import numpy as np
import pytest
import torch

from tests.cases import case_es, hip_solver, oracle_solver

pytestmark = pytest.mark.gpu

CASES = {
    "1d1v": dict(),
    "2d2v": dict(N_q1=16, N_q2=12, N_p=10),
    "1d1v_weno": dict(recon_q="weno5", recon_p="weno5", N_p=20),
    "1d1v_B": dict(fields_initialize="fft + user-defined magnetic fields", B0=0.3),
}


def _interior(x, g1, g2, n1, n2):
    return x[:, :, g1:g1 + n1, g2:g2 + n2].detach().cpu().double().numpy()


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("stepper", ["RK2", "SSP-RK3"])
def test_electrostatic_per_step(gpu, name, stepper):
    c = case_es(seed=4, stepper=stepper, **CASES[name])
    o, s = oracle_solver(c), hip_solver(c)
    g1, g2, n1, n2 = s.N_g1, s.N_g2, s.N_q1, s.N_q2
    fo = _interior(o.fields_solver.cell_centered_EM_fields, g1, g2, n1, n2)
    fh = _interior(s.fields_solver.cell_centered_EM_fields, g1, g2, n1, n2)
    assert np.abs(fh - fo).max() <= 1e-13 * max(np.abs(fo).max(), 1e-30)
    worst = 0.0
    for _ in range(3):
        s.f = o.f.to(s.device)
        o.strang_timestep(c.dt)
        s.strang_timestep(c.dt)
        a = _interior(s.f, g1, g2, n1, n2)
        b = o.interior().numpy()
        worst = max(worst, float(np.abs(a - b).max() / np.abs(b).max()))
    print(f"{name} {stepper}: per-step rel Linf {worst:.3e}", flush=True)
    assert worst <= 1e-12
