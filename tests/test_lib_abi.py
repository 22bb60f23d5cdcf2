"""The C-ABI library loads and exports every entry point include/boltx.h
declares; argument checks reject bad input before any device work."""
import ctypes
import os
import re

from bolt_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "boltx.h")).read()
    return sorted(set(re.findall(r"\b(boltx_[a-z_]+)\s*\(", text)))


def test_exports_match_header():
    lib = _lib.load()
    names = declared()
    assert set(names) == set(_lib.EXPORTS), names
    for n in names:
        assert hasattr(lib, n)
    assert lib.boltx_version() == 1


def test_invalid_arguments_are_rejected():
    lib = _lib.load()
    rc = lib.boltx_fvm_stage(None, None, None, None)
    assert rc == -1
    assert b"grid" in lib.boltx_last_error()
    g = _lib.Grid(4, 4, 8, 8, 1, 2, 0, 0.25, 0.25, 1.0, 1.0, 1.0, -4.0, -4.0, -0.5)
    ph = _lib.Physics(3, 3, 0, 0, 1, 0, 2, 7, 1.0, 1.0, 1.0, 0.01)
    rc = lib.boltx_moments(ctypes.byref(g), ctypes.byref(ph), ctypes.c_void_p(1),
                           ctypes.c_void_p(1), None, None)
    assert rc == -3 and b"species" in lib.boltx_last_error()
    g.n_species = 1
    st = _lib.Stage()
    st.u = st.w = st.out = ctypes.c_void_p(8)
    rc = lib.boltx_fvm_stage(ctypes.byref(g), ctypes.byref(ph), ctypes.byref(st), None)
    assert rc == -1 and b"precision" in lib.boltx_last_error()
    ph.precision = 0
    rc = lib.boltx_fvm_stage(ctypes.byref(g), ctypes.byref(ph), ctypes.byref(st), None)
    assert rc == -1 and b"alias" in lib.boltx_last_error()
    # the slab-ordered kernel needs a workspace of two int tables per slab lane
    g2 = _lib.Grid(16, 16, 48, 48, 1, 1, 0, 1 / 16, 1 / 16, 20 / 48, 20 / 48, 1.0, -10.0, -10.0, -0.5)
    assert lib.boltx_workspace_size(ctypes.byref(g2), ctypes.byref(ph)) == 2 * 36 * 64 * 4  # 36 slabs of 64 lanes
    g2.n_p1 = 2000
    assert lib.boltx_workspace_size(ctypes.byref(g2), ctypes.byref(ph)) == 0
    # the q1-line layout needs tables (n_p1+n_p2+n_p3 rows of Q) and moment partials
    g2.n_p1, g2.layout = 48, 1
    tab = (48 + 48 + 1) * 256 * 8
    assert lib.boltx_workspace_size(ctypes.byref(g2), ctypes.byref(ph)) == \
        (tab + 255) // 256 * 256 + 8 * 5 * 256 * 8  # 8 p-chunks of moment partials
    g2.n_q1 = 512
    rc = lib.boltx_moments(ctypes.byref(g2), ctypes.byref(ph), ctypes.c_void_p(1),
                           ctypes.c_void_p(1), ctypes.c_void_p(1), None)
    assert rc == -3 and b"256" in lib.boltx_last_error()


def test_adjoint_and_ensemble_reject_bad_input():
    """Argument checks of the adjoint and ensemble entry points, before any
    device work."""
    import ctypes as C
    lib = _lib.load()
    g = _lib.Grid(8, 8, 6, 6, 1, 1, 0, 0.125, 0.125, 1.0, 1.0, 1.0, -3.0, -3.0, -0.5)
    ph = _lib.Physics(3, 3, 1, 1, 1, 0, 2, 1, 1.0, -1.0, 1.0, 0.01)  # FP32-bulk
    adj = _lib.Adjoint(C.c_void_p(16), None, C.c_void_p(32), None, C.c_void_p(48),
                       0.0, 1.0, 1.0, None, None)
    assert lib.boltx_adjoint_stage(C.byref(g), C.byref(ph), C.byref(adj), None) == -3
    assert b"FP64" in lib.boltx_last_error()
    ph.precision = 0
    assert lib.boltx_adjoint_stage(C.byref(g), C.byref(ph), C.byref(adj), None) == -1
    assert b"mom_u" in lib.boltx_last_error()
    adj.out = adj.lam
    assert lib.boltx_adjoint_stage(C.byref(g), C.byref(ph), C.byref(adj), None) == -1
    assert b"alias" in lib.boltx_last_error()
    adj.out, adj.mom_u, adj.workspace = C.c_void_p(48), C.c_void_p(64), C.c_void_p(80)
    g.n_ghost, g.bc[0], g.bc[1] = 3, _lib.BC["mirror"], _lib.BC["mirror"]
    assert lib.boltx_adjoint_stage(C.byref(g), C.byref(ph), C.byref(adj), None) == -3
    assert b"periodic" in lib.boltx_last_error()
    g.n_ghost, g.bc[0], g.bc[1] = 0, 0, 0
    assert lib.boltx_adjoint_workspace_size(C.byref(g), C.byref(ph)) == (5 + 6 * 8) * 64 * 8
    m = (_lib.EnsMember * 1)()
    m[0].grid = g
    m[0].grid.n_q2 = 1
    m[0].phys = ph
    m[0].dt = 0.0
    t = _lib.EnsTuning(16384, 2048)
    assert lib.boltx_ensemble_step(1, m, C.byref(t), 1, None, None, None, None, 1, None) == -1
    assert b"dt" in lib.boltx_last_error()


def test_adjoint_rejects_large_grid_before_launching():
    """The size limits of the adjoint's launch grid are checked before its
    first kernel: on a machine without a GPU the call returns the size error
    (-3), not a launch failure (-2)."""
    import ctypes as C
    lib = _lib.load()
    g = _lib.Grid(256, 8, 256, 256, 128, 1, 1, 1 / 256, 1 / 8, 0.1, 0.1, 0.1, -10.0, -10.0,
                  -6.0)
    ph = _lib.Physics(3, 3, 1, 1, 1, 0, 3, 0, 1.0, -1.0, 1.0, 0.01)
    adj = _lib.Adjoint(C.c_void_p(16), C.c_void_p(24), C.c_void_p(32), None, C.c_void_p(48),
                       0.0, 1.0, 1.0, None, C.c_void_p(64))
    assert lib.boltx_adjoint_stage(C.byref(g), C.byref(ph), C.byref(adj), None) == -3
    assert b"2^31" in lib.boltx_last_error()



def test_noalloc_stage_rejects_before_device_work():
    """boltx_fvm_stage_noalloc refuses a caller workspace and a problem whose
    workspace exceeds its scratch array, before touching the device (on a
    machine without a GPU the calls return these errors, not a launch
    failure)."""
    import ctypes as C
    lib = _lib.load()
    g = _lib.Grid(256, 4096, 48, 48, 1, 1, 1, 1 / 256, 1 / 4096, 20 / 48, 20 / 48, 1.0,
                  -10.0, -10.0, -0.5)
    ph = _lib.Physics(3, 3, 1, 1, 1, 0, 2, 0, 1.0, -1.0, 1.0, 0.01)
    assert lib.boltx_workspace_size(C.byref(g), C.byref(ph)) > _lib.SCRATCH_BYTES
    st = _lib.Stage()
    st.u, st.w, st.out, st.mom_in = C.c_void_p(8), C.c_void_p(8), C.c_void_p(16), C.c_void_p(24)
    st.a, st.b, st.cdt = 0.0, 1.0, 1e-3
    assert lib.boltx_fvm_stage_noalloc(C.byref(g), C.byref(ph), C.byref(st), None) == -3
    assert b"scratch" in lib.boltx_last_error()
    st.workspace = C.c_void_p(32)
    assert lib.boltx_fvm_stage_noalloc(C.byref(g), C.byref(ph), C.byref(st), None) == -1
    assert b"NULL" in lib.boltx_last_error()
