"""Drive the CPU harness (libstage_host.so) with the oracle's state."""
import ctypes
import os
import sys

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "bolt_amd"))
import _lib as L  # noqa: E402  (ctypes structs only; no library load)

from . import build_host  # noqa: E402

_libs = {}


def lib(contract="fast-honor-pragmas"):
    if contract not in _libs:
        path = build_host.build(contract)
        h = ctypes.CDLL(path)
        h.host_fvm_stage.argtypes = [ctypes.POINTER(L.Grid), ctypes.POINTER(L.Physics),
                                     ctypes.POINTER(L.Stage)]
        h.host_fvm_stage.restype = ctypes.c_int
        h.host_fvm_stage_coeffs.argtypes = [ctypes.POINTER(L.Grid), ctypes.POINTER(L.Physics),
                                            ctypes.POINTER(L.Stage), ctypes.POINTER(L.Coeffs)]
        h.host_fvm_stage_coeffs.restype = ctypes.c_int
        _libs[contract] = h
    return _libs[contract]


def to_internal(o, x):
    """oracle (P, 1, Q1g, Q2g) -> numpy [N_q2][N_q1][P] interior."""
    return np.ascontiguousarray(o.interior(x)[:, 0].permute(2, 1, 0).numpy())


def grid_physics(o, precision=0, source=None, padv=None):
    p = o.params
    g = L.Grid(o.N_q1, o.N_q2, o.N_p1, o.N_p2, o.N_p3, 1, 0, o.dq1, o.dq2,
               float(o.dp1), float(o.dp2), float(o.dp3),
               float(o.physical_system.p1_start[0]), float(o.physical_system.p2_start[0]),
               float(o.physical_system.p3_start[0]))
    tau = float(p.tau(o.q1_center, o.q2_center, o.p1_center, o.p2_center,
                      o.p3_center).reshape(-1)[0])
    ph = L.Physics(L.RECON[p.reconstruction_method_in_q], L.RECON[p.reconstruction_method_in_p],
                   L.RIEMANN[p.riemann_solver_in_q], L.RIEMANN[p.riemann_solver_in_p],
                   int(p.source_enabled if source is None else source),
                   int((p.fields_enabled and p.solver_method_in_p == "FVM") if padv is None else padv),
                   int(p.p_dim), precision, float(o.mass.reshape(-1)[0]),
                   float(o.charge.reshape(-1)[0]), float(p.boltzmann_constant), tau)
    return g, ph


def moments_np(o, u_int):
    """Five built-in moments of an interior array, FP64 [5][Q] (oracle-defined)."""
    P = o.N_p1 * o.N_p2 * o.N_p3
    v1 = o.p1_center.reshape(P).numpy()
    v2 = o.p2_center.reshape(P).numpy()
    v3 = o.p3_center.reshape(P).numpy()
    m = float((o.dp3 * o.dp2 * o.dp1).reshape(-1)[0])
    f = u_int.reshape(-1, P).astype(np.float64)
    return np.ascontiguousarray(np.stack([
        f.sum(1) * m, (f * v1).sum(1) * m, (f * v2).sum(1) * m, (f * v3).sum(1) * m,
        (0.5 * f * (v1**2 + v2**2 + v3**2)).sum(1) * m]))


def stage(o, u, w, un, a, b, cdt, precision=0, contract="fast-honor-pragmas", fields=None, **kw):
    """Run one harness stage on interior numpy arrays; returns (out, mom_out)."""
    g, ph = grid_physics(o, precision, **kw)
    dt = np.float64 if precision == 0 else np.float32
    u = np.ascontiguousarray(u, dtype=dt)
    w = np.ascontiguousarray(w, dtype=dt)
    un = None if un is None else np.ascontiguousarray(un, dtype=dt)
    out = np.empty_like(u)
    mom_in = moments_np(o, u)
    mom_out = np.empty_like(mom_in)
    P = lambda x: None if x is None else x.ctypes.data_as(ctypes.c_void_p)
    F = [None, None, None] if fields is None else [np.ascontiguousarray(x) for x in fields]
    st = L.Stage(P(u), P(w), P(un), P(out), a, b, cdt, P(mom_in), P(mom_out),
                 P(F[0]), P(F[1]), P(F[2]), None)
    lib(contract).host_fvm_stage(ctypes.byref(g), ctypes.byref(ph), ctypes.byref(st))
    return out, mom_out


def coeff(o, x, keep):
    """boltx_coeff of a callback result ``x`` broadcastable to the oracle's
    (P, 1, Q1g, Q2g), origin at the first interior q-cell (the layout the
    solver hands to boltx_fvm_stage_coeffs)."""
    P = o.N_p1 * o.N_p2 * o.N_p3
    x = torch.as_tensor(x, dtype=torch.float64)
    v = torch.broadcast_to(x, (P, 1) + tuple(o.q1_center.shape[2:]))[:, 0]
    keep.append(v)
    off = v.stride(1) * o.N_g1 + v.stride(2) * o.N_g2
    return L.Coeff(v.data_ptr() + 8 * off, v.stride(0), v.stride(1), v.stride(2))


def stage_coeffs(o, u, w, un, a, b, cdt, co, precision=0, fields=None, **kw):
    """One harness stage with caller-supplied terms ``co`` (boltx_coeffs)."""
    g, ph = grid_physics(o, precision, **kw)
    dt = np.float64 if precision == 0 else np.float32
    u = np.ascontiguousarray(u, dtype=dt)
    w = np.ascontiguousarray(w, dtype=dt)
    un = None if un is None else np.ascontiguousarray(un, dtype=dt)
    out = np.empty_like(u)
    mom_in = moments_np(o, u)
    mom_out = np.empty_like(mom_in)
    P = lambda x: None if x is None else x.ctypes.data_as(ctypes.c_void_p)
    F = [None, None, None] if fields is None else [np.ascontiguousarray(x) for x in fields]
    st = L.Stage(P(u), P(w), P(un), P(out), a, b, cdt, P(mom_in), P(mom_out),
                 P(F[0]), P(F[1]), P(F[2]), None)
    lib().host_fvm_stage_coeffs(ctypes.byref(g), ctypes.byref(ph), ctypes.byref(st),
                                ctypes.byref(co))
    return out, mom_out
