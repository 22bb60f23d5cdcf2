"""GPU parity of multi-species problems (S = 2: own mass, charge, p-grid,
temperature and collision time per species) against the FP64 oracle, with
the fields off (both storage layouts), electrostatic and electrodynamic."""
import os
import tempfile

import numpy as np
import pytest
import torch

from tests.cases import case_species, hip_solver, oracle_solver
from tests.test_gpu_parity import _cmp, _per_step

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("layout,n1", [("reference", 16), ("q1-lines", 64)])
@pytest.mark.parametrize("stepper", ["RK2", "SSP-RK3"])
def test_two_species_fields_off(gpu, layout, n1, stepper):
    c = case_species(N_q1=n1, memory_layout=layout, stepper=stepper, recon_q="weno5")
    nls = hip_solver(c)
    assert nls._layout == (0 if layout == "reference" else 1)
    print(f"two species, fields off, {layout}, {stepper}", flush=True)
    _per_step(c, 3, 1e-12)


def test_two_species_electrostatic(gpu):
    """Ions (m = 25, q = +1, p in [-1.5, 1.5]) and electrons (m = 1, q = -1,
    p in [-6, 6]) in 1D1V with the FFT Poisson field from the summed rho."""
    c = case_species(masses=(1.0, 25.0), charges=(-1.0, 1.0), temps=(1.0, 1.0),
                     drifts=(0.0, 0.0), p_max=(6.0, 1.5), taus=(0.1, 0.5), N_q1=32, N_q2=1,
                     N_p1=16, N_p2=1, p_dim=1, fields=True, fields_type="electrostatic",
                     recon_q="minmod", recon_p="minmod", riemann="lax-friedrichs", amp=5e-2)
    _per_step(c, 4, 1e-12)
    o, nls = oracle_solver(c), hip_solver(c)
    for _ in range(20):
        o.strang_timestep(c.dt)
        nls.strang_timestep(c.dt)
    linf, _ = _cmp(nls, o)
    print(f"two species electrostatic 20 steps: rel Linf {linf:.3e}", flush=True)
    assert linf <= 1e-11
    E = nls.fields_solver.cell_centered_EM_fields.cpu().numpy()
    Eo = o.fields_solver.cell_centered_EM_fields.numpy()
    g = nls.N_g1
    d = np.abs(E[:2, :, g:-g] - Eo[:2, :, g:-g]).max() / np.abs(Eo[:2, :, g:-g]).max()
    assert d <= 1e-11


def test_two_species_electrodynamic(gpu):
    """2D3V with the FDTD fields driven by the current summed over species."""
    c = case_species(N_q1=8, N_q2=6, N_p1=6, N_p2=6, N_p3=4, p_dim=3, p_max=(5.0, 2.5),
                     fields=True, fields_type="electrodynamic", recon_q="ppm",
                     recon_p="weno5", riemann="lax-friedrichs", stepper="RK2", dt=5e-4)
    _per_step(c, 3, 1e-12)


def test_two_species_moments_and_fp32_bulk(gpu):
    c = case_species(stepper="SSP-RK3")
    o, nls = oracle_solver(c), hip_solver(c)
    for _ in range(5):
        o.strang_timestep(c.dt)
        nls.strang_timestep(c.dt)
    g = nls.N_ghost
    for name in ("density", "mom_v1_bulk", "energy"):
        a = nls.compute_moments(name).cpu().numpy()[..., g:g + 16, g:g + 6]
        b = o.compute_moments(name).numpy()[..., g:g + 16, g:g + 6]
        assert a.shape == b.shape == (1, 2, 16, 6)
        assert float(np.abs(a - b).max() / np.abs(b).max()) <= 1e-11, name
    c32 = case_species(stepper="SSP-RK3", precision="fp32-bulk")
    n32 = hip_solver(c32)
    for _ in range(5):
        n32.strang_timestep(c32.dt)
    x, y = nls.f.double(), n32.f.double()
    rel = float(torch.linalg.vector_norm(x - y) / torch.linalg.vector_norm(x))
    print(f"two species fp32-bulk vs fp64: rel L2 {rel:.3e}", flush=True)
    assert rel <= 1e-5


def test_two_species_checkpoint_round_trip(gpu):
    """dump_distribution_function / load_distribution_function with S = 2
    (PETSc natural order, dof = p + P s) restore the state bit for bit."""
    c = case_species(stepper="RK2")
    nls = hip_solver(c)
    for _ in range(3):
        nls.strang_timestep(c.dt)
    before = nls.f.clone()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "f2")
        nls.dump_distribution_function(path)
        n = os.path.getsize(path + ".bin")
        assert n == 8 + 8 * nls._Q * nls._P * 2  # classid, N, then N doubles
        other = hip_solver(c)
        other.load_distribution_function(path)
    assert torch.equal(other.f, before)
