"""boltx_fvm_stage_noalloc (the stage with the library's static scratch array
as its workspace) against boltx_fvm_stage with a caller-owned workspace.

Both run the same kernels, so outputs and moment blocks must agree bit for
bit.  Each layout is checked on two no-allocation calls in a row, the second
with a smaller problem than the first, so the second runs over what the first
left in the scratch array."""
import ctypes

import pytest
import torch

from bolt_amd import _lib
from tests.cases import case_2d2v, case_2d3v, hip_solver

pytestmark = pytest.mark.gpu


def _stage_inputs(nls, seed):
    """A stage of the SSP-RK3 form out = 3/4 un + 1/4 (u + dt L(u)) over the
    solver's initial state: u, un = u perturbed, random cell-centred fields
    for C_p (used when p-advection is on)."""
    g = torch.Generator(device=nls.device).manual_seed(seed)
    u = nls._f.clone()
    noise = torch.rand(u.shape, generator=g, device=nls.device, dtype=torch.float64)
    un = (u.double() * (1 + 1e-3 * noise)).to(u.dtype)
    F = [0.1 * torch.randn(6, nls._Q, generator=g, device=nls.device, dtype=torch.float64)
         for _ in range(3)]
    return u, un, nls._mom.clone(), F


def _call(nls, entry, inputs, workspace):
    u, un, mom_in, F = inputs
    out = torch.full_like(u, float("nan"))
    mom_out = torch.full_like(mom_in, float("nan"))
    st = _lib.Stage(_lib.ptr(u), _lib.ptr(u), _lib.ptr(un), _lib.ptr(out), 0.75, 0.25, 5e-4,
                    _lib.ptr(mom_in), _lib.ptr(mom_out), _lib.ptr(F[0]), _lib.ptr(F[1]),
                    _lib.ptr(F[2]), _lib.ptr(workspace))
    _lib.check(getattr(nls._L, entry)(ctypes.byref(nls._grid), ctypes.byref(nls._phys),
                                      ctypes.byref(st), _lib.stream_handle()))
    return out, mom_out


def _check_pair(big, small):
    """Reference results first, then the two no-allocation calls back to back."""
    runs = []
    for k, nls in enumerate((big, small)):
        inputs = _stage_inputs(nls, seed=10 + k)
        ws = nls._ws
        assert ws is None or ws.numel() <= _lib.SCRATCH_BYTES
        runs.append((nls, inputs, _call(nls, "boltx_fvm_stage", inputs, ws)))
    got = [_call(nls, "boltx_fvm_stage_noalloc", inputs, None) for nls, inputs, _ in runs]
    torch.cuda.synchronize()
    for (nls, _, (out_ref, mom_ref)), (out, mom) in zip(runs, got):
        assert torch.isfinite(out_ref).all()
        assert torch.equal(out, out_ref), (nls._fshape, (out - out_ref).abs().max().item())
        assert torch.equal(mom, mom_ref)


@pytest.mark.parametrize("precision", ["fp64", "fp32-bulk"])
def test_noalloc_q1_lines(gpu, precision):
    """Fields off, q1-line layout: the scratch holds the Maxwellian tables and
    the moment partials."""
    big = hip_solver(case_2d2v(seed=3, N_q=128, N_p=16, precision=precision))
    small = hip_solver(case_2d2v(seed=4, N_q=64, N_p=12, precision=precision))
    assert big._layout == small._layout == _lib.LAYOUT["q1-lines"]
    print(f"no-alloc q1-lines {precision}: {big._fshape} then {small._fshape}", flush=True)
    _check_pair(big, small)


@pytest.mark.parametrize("precision", ["fp64", "fp32-bulk"])
def test_noalloc_reference_layout(gpu, precision):
    """Fields on (p-advection), reference layout: the scratch holds the
    sign-grouped p order, rewritten by each call."""
    big = hip_solver(case_2d3v(seed=5, N_q=12, N_p=10, precision=precision))
    small = hip_solver(case_2d3v(seed=6, N_q=8, N_p=6, precision=precision))
    assert big._layout == small._layout == _lib.LAYOUT["reference"]
    assert big._phys.p_advection == 1
    print(f"no-alloc reference {precision}: {big._fshape} then {small._fshape}", flush=True)
    _check_pair(big, small)
