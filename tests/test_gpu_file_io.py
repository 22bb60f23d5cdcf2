"""Checkpoint and restart through the reference's dump/load methods
(nonlinear_solver.py:513-519) in PETSc binary format, on the GPU solver."""
import numpy as np
import pytest
import torch

from bolt_amd import file_io
from tests.cases import case_2d2v, case_2d3v, hip_solver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("layout", ["reference", "q1-lines"])
def test_restart_continues_the_run(gpu, tmp_path, layout):
    """dump after 5 steps, load into a fresh solver: the next 5 steps of both
    are bit-identical (the moment blocks and the time come back with f)."""
    c = case_2d2v(seed=2, N_q=64 if layout == "q1-lines" else 12, N_p=8, memory_layout=layout)
    a = hip_solver(c)
    for _ in range(5):
        a.strang_timestep(c.dt)
    name = str(tmp_path / "dist")
    a.dump_distribution_function(name)
    b = hip_solver(c)
    b.load_distribution_function(name)
    assert torch.equal(a._f, b._f) and torch.equal(a._mom, b._mom)
    assert b.time_elapsed == a.time_elapsed
    for _ in range(5):
        a.strang_timestep(c.dt)
        b.strang_timestep(c.dt)
    assert torch.equal(a._f, b._f)
    assert b.time_elapsed == a.time_elapsed
    # the file holds the interior in natural order: dof (p) fastest
    g1, g2 = b.N_g1, b.N_g2
    f_ref = b.f.cpu().numpy()
    b.dump_distribution_function(name)
    v = file_io.read_vec(name + ".bin")
    want = f_ref[:, :, g1:g1 + b.N_q1, g2:g2 + b.N_q2].reshape(-1, order="F")
    assert np.array_equal(v, want)


def test_restart_from_reference_files_recomputes(gpu, tmp_path):
    """Without the companion state (a file the reference wrote, or a .bin
    replaced after the dump) the load recomputes the moments from f, as the
    reference's load does; the run then continues to within the rounding
    that the changed summation order starts (amplified by WENO5's expanded
    indicator on smooth data)."""
    c = case_2d2v(seed=2, N_q=12, N_p=8, memory_layout="reference")
    a = hip_solver(c)
    for _ in range(5):
        a.strang_timestep(c.dt)
    name = str(tmp_path / "dist")
    a.dump_distribution_function(name)
    # a stale companion (the .bin changed after the dump) is not used
    raw = bytearray(open(name + ".bin", "rb").read())
    raw[-1] ^= 1
    open(name + ".bin", "wb").write(bytes(raw))
    b = hip_solver(c)
    b.load_distribution_function(name)
    assert b.time_elapsed == 0
    # the reference's own files: no companion state at all
    import os
    a.dump_distribution_function(name)
    os.remove(name + ".state.json")
    b = hip_solver(c)
    b.load_distribution_function(name)
    assert torch.equal(a._f, b._f)
    for _ in range(5):
        a.strang_timestep(c.dt)
        b.strang_timestep(c.dt)
    d = float((a._f - b._f).abs().max() / a._f.abs().max())
    assert d <= 1e-11, d


@pytest.mark.parametrize("stepper", ["RK2", "SSP-RK3"])
def test_em_restart_bit_identical(gpu, tmp_path, stepper):
    """Fields on (FDTD; RK2 with the reference's toggle, SSP-RK3 once per
    step): f and the fields dumped mid-run and loaded into a fresh solver
    continue bit for bit (the staggered time levels and at_n come back)."""
    c = case_2d3v(seed=4, N_q=8, N_p=6, stepper=stepper)
    a = hip_solver(c)
    for _ in range(3):
        a.strang_timestep(c.dt)
    fn, en = str(tmp_path / "f"), str(tmp_path / "em")
    a.dump_distribution_function(fn)
    a.dump_EM_fields(en)
    b = hip_solver(c)
    b.load_distribution_function(fn)
    b.load_EM_fields(en)
    fa, fb = a.fields_solver, b.fields_solver
    assert fb.at_n == fa.at_n and fb.time_elapsed == fa.time_elapsed
    assert (fb._cc_at_n is fb._cc) == (fa._cc_at_n is fa._cc)
    for _ in range(3):
        a.strang_timestep(c.dt)
        b.strang_timestep(c.dt)
    assert torch.equal(a._f, b._f)
    assert torch.equal(fa._yee, fb._yee)
    for x, y in ((fa._cc, fb._cc), (fa._cc_at_n, fb._cc_at_n),
                 (fa._cc_at_n_plus_half, fb._cc_at_n_plus_half)):
        assert torch.equal(x, y)
    assert b.time_elapsed == a.time_elapsed


def test_moments_dump_holds_every_moment(gpu, tmp_path):
    c = case_2d2v(seed=1, N_q=12, N_p=8)
    s = hip_solver(c)
    s.strang_timestep(c.dt)
    name = str(tmp_path / "mom")
    s.dump_moments(name)
    names = file_io.moment_names(s)
    v = file_io.read_vec(name + ".bin").reshape(s.N_q2, s.N_q1, len(names))
    g1, g2 = s.N_g1, s.N_g2
    for k, n in enumerate(names):
        m = s.compute_moments(n)[0, 0, g1:g1 + s.N_q1, g2:g2 + s.N_q2].cpu().numpy()
        assert np.array_equal(v[:, :, k], m.T), n


def test_em_fields_round_trip(gpu, tmp_path):
    c = case_2d3v(seed=3, N_q=8, N_p=6)
    s = hip_solver(c)
    for _ in range(2):
        s.strang_timestep(c.dt)
    name = str(tmp_path / "em")
    s.dump_EM_fields(name)
    yee = s.fields_solver.yee_grid_EM_fields.clone()
    t = hip_solver(c)
    t.load_EM_fields(name)
    assert torch.equal(t.fields_solver.yee_grid_EM_fields, yee)
    # with the companion state: the solver's own cell-centred fields (after a
    # step its E rows hold E at n+1/2: fields.py:640)
    assert torch.equal(t.fields_solver._cc, s.fields_solver._cc)
    # without it, load.py:76: the cell-centred fields of the loaded Yee grid
    import os
    os.remove(name + ".state.json")
    cc = torch.empty_like(s.fields_solver._cc)
    s.fields_solver._yee_to_cc(cc)
    t = hip_solver(c)
    t.load_EM_fields(name)
    assert torch.equal(t.fields_solver._cc, cc)
