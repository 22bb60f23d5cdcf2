"""ctypes binding of ``libboltx_hip.so`` (declarations in include/boltx.h).

Loading fails loudly: there is no fallback path.  Every call passes device
pointers from torch tensors and the current torch stream.
"""
import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libboltx_hip.so")

RECON = {"piecewise-constant": 0, "minmod": 1, "ppm": 2, "weno5": 3}
RIEMANN = {"upwind-flux": 0, "lax-friedrichs": 1}
PRECISION = {"fp64": 0, "fp32-bulk": 1}
LAYOUT = {"reference": 0, "q1-lines": 1}
BC = {"periodic": 0, "none": 0, "mirror": 1, "dirichlet": 2, "mirror+dirichlet": 3}
STEPPER = {"RK2": 0, "SSP-RK3": 1}

EXPORTS = ("boltx_fvm_stage", "boltx_fvm_stage_noalloc", "boltx_fvm_stage_coeffs",
           "boltx_moments",
           "boltx_workspace_size", "boltx_workspace_init",
           "boltx_currents", "boltx_fdtd_step", "boltx_yee_to_cell_centered",
           "boltx_check_finite", "boltx_adjoint_stage", "boltx_adjoint_workspace_size",
           "boltx_ensemble_layout", "boltx_ensemble_workspace_size", "boltx_ensemble_map",
           "boltx_ensemble_init", "boltx_ensemble_step", "boltx_fill_ghosts",
           "boltx_version", "boltx_last_error")
MOM_ROWS = 12
SCRATCH_BYTES = 64 << 20  # BOLTX_SCRATCH_BYTES


class Grid(ctypes.Structure):
    _fields_ = [("n_q1", ctypes.c_int32), ("n_q2", ctypes.c_int32),
                ("n_p1", ctypes.c_int32), ("n_p2", ctypes.c_int32), ("n_p3", ctypes.c_int32),
                ("n_species", ctypes.c_int32), ("layout", ctypes.c_int32),
                ("dq1", ctypes.c_double), ("dq2", ctypes.c_double),
                ("dp1", ctypes.c_double), ("dp2", ctypes.c_double), ("dp3", ctypes.c_double),
                ("p1_start", ctypes.c_double), ("p2_start", ctypes.c_double),
                ("p3_start", ctypes.c_double), ("n_ghost", ctypes.c_int32),
                ("bc", ctypes.c_int32 * 4)]


class Physics(ctypes.Structure):
    _fields_ = [("recon_q", ctypes.c_int32), ("recon_p", ctypes.c_int32),
                ("riemann_q", ctypes.c_int32), ("riemann_p", ctypes.c_int32),
                ("source_enabled", ctypes.c_int32), ("p_advection", ctypes.c_int32),
                ("p_dim", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("mass", ctypes.c_double), ("charge", ctypes.c_double),
                ("boltzmann_constant", ctypes.c_double), ("tau", ctypes.c_double)]


class Stage(ctypes.Structure):
    _fields_ = [("u", ctypes.c_void_p), ("w", ctypes.c_void_p), ("un", ctypes.c_void_p),
                ("out", ctypes.c_void_p),
                ("a", ctypes.c_double), ("b", ctypes.c_double), ("cdt", ctypes.c_double),
                ("mom_in", ctypes.c_void_p), ("mom_out", ctypes.c_void_p),
                ("fields_c_p1", ctypes.c_void_p), ("fields_c_p2", ctypes.c_void_p),
                ("fields_c_p3", ctypes.c_void_p), ("workspace", ctypes.c_void_p),
                ("currents_out", ctypes.c_void_p)]


class Coeff(ctypes.Structure):
    _fields_ = [("ptr", ctypes.c_void_p), ("stride_p", ctypes.c_int64),
                ("stride_q1", ctypes.c_int64), ("stride_q2", ctypes.c_int64)]


class Coeffs(ctypes.Structure):
    _fields_ = [("c_q1", Coeff), ("c_q2", Coeff), ("c_p1", Coeff), ("c_p2", Coeff),
                ("c_p3", Coeff), ("source", Coeff), ("tau", Coeff)]


class Adjoint(ctypes.Structure):
    _fields_ = [("u", ctypes.c_void_p), ("mom_u", ctypes.c_void_p), ("lam", ctypes.c_void_p),
                ("base", ctypes.c_void_p), ("out", ctypes.c_void_p),
                ("a", ctypes.c_double), ("b", ctypes.c_double), ("c", ctypes.c_double),
                ("tau_grad", ctypes.c_void_p), ("workspace", ctypes.c_void_p),
                ("base2", ctypes.c_void_p), ("a2", ctypes.c_double)]


class EnsMember(ctypes.Structure):
    _fields_ = [("grid", Grid), ("phys", Physics), ("dt", ctypes.c_double)]


class EnsTuning(ctypes.Structure):
    _fields_ = [("chunk_cells", ctypes.c_int32), ("batch_cells", ctypes.c_int32)]


class BoltxError(RuntimeError):
    pass


_lib = None


def load(path=LIB_PATH):
    """Load the HIP library (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise BoltxError(f"{path} is missing: build it with `python -m bolt_amd.build` "
                         "(there is no fallback path)")
    lib = ctypes.CDLL(path)
    for name in EXPORTS:
        getattr(lib, name)
    vp, i32, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_double
    g, ph, st = ctypes.POINTER(Grid), ctypes.POINTER(Physics), ctypes.POINTER(Stage)
    lib.boltx_fvm_stage.argtypes = [g, ph, st, vp]
    lib.boltx_fvm_stage_noalloc.argtypes = [g, ph, st, vp]
    lib.boltx_fvm_stage_coeffs.argtypes = [g, ph, st, ctypes.POINTER(Coeffs), vp]
    lib.boltx_moments.argtypes = [g, ph, vp, vp, vp, vp]
    lib.boltx_workspace_size.argtypes = [g, ph]
    lib.boltx_workspace_init.argtypes = [g, ph, vp, vp]
    lib.boltx_currents.argtypes = [g, ph, vp, vp, vp]
    lib.boltx_fdtd_step.argtypes = [g, vp, vp, dbl, dbl, dbl, vp]
    lib.boltx_yee_to_cell_centered.argtypes = [g, vp, vp, vp]
    lib.boltx_check_finite.argtypes = [g, vp, vp, vp]
    lib.boltx_adjoint_stage.argtypes = [g, ph, ctypes.POINTER(Adjoint), vp]
    lib.boltx_adjoint_workspace_size.argtypes = [g, ph]
    em, tu, i32 = ctypes.POINTER(EnsMember), ctypes.POINTER(EnsTuning), ctypes.c_int
    lib.boltx_ensemble_layout.argtypes = [i32, em, ctypes.POINTER(ctypes.c_longlong)]
    lib.boltx_ensemble_workspace_size.argtypes = [i32, em, tu]
    lib.boltx_ensemble_map.argtypes = [i32, em, tu, ctypes.POINTER(ctypes.c_size_t),
                                       ctypes.POINTER(ctypes.c_size_t)]
    lib.boltx_ensemble_init.argtypes = [i32, em, tu, vp, vp, vp]
    lib.boltx_ensemble_step.argtypes = [i32, em, tu, i32, vp, vp, vp, vp, i32, vp]
    lib.boltx_fill_ghosts.argtypes = [g, ph, vp, ctypes.POINTER(vp), vp]
    for name in EXPORTS[:-2]:
        getattr(lib, name).restype = ctypes.c_int
    lib.boltx_workspace_size.restype = ctypes.c_size_t
    lib.boltx_adjoint_workspace_size.restype = ctypes.c_size_t
    lib.boltx_ensemble_layout.restype = ctypes.c_longlong
    lib.boltx_ensemble_workspace_size.restype = ctypes.c_size_t
    lib.boltx_version.argtypes = []
    lib.boltx_version.restype = ctypes.c_int
    lib.boltx_last_error.argtypes = []
    lib.boltx_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        msg = _lib.boltx_last_error().decode() if _lib is not None else ""
        raise BoltxError(f"boltx call failed ({rc}): {msg}")


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_handle():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
