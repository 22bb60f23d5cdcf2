"""Ensemble planning through the C-ABI on the host (no device work): pooled
layout offsets, workspace sizing and the argument checks."""
import ctypes

from bolt_amd import _lib


def _member(n1, npp, dt=1e-3, prec=1, recon=3):
    m = _lib.EnsMember()
    dp = 20.0 / npp
    m.grid = _lib.Grid(n1, 1, npp, npp, 1, 1, 1, 1.0 / n1, 1.0, dp, dp, 1.0, -10.0, -10.0, -0.5)
    m.phys = _lib.Physics(recon, recon, 1, 1, 1, 0, 2, prec, 1.0, -1.0, 1.0, 0.01)
    m.dt = dt
    return m


def test_layout_offsets_pool_members_in_order():
    L = _lib.load()
    sizes = [(32, 24), (256, 64), (64, 32)]
    arr = (_lib.EnsMember * 3)(*[_member(a, b) for a, b in sizes])
    offs = (ctypes.c_longlong * 3)()
    total = L.boltx_ensemble_layout(3, arr, offs)
    want, o = [], 0
    for n1, npp in sizes:
        want.append(o)
        o += n1 * npp * npp
    assert list(offs) == want and total == o


def test_workspace_grows_with_chunks_and_checks_input():
    """Smaller chunks mean more partial-moment blocks in the workspace; the
    batch size does not change the workspace; non-positive sizes are
    rejected."""
    L = _lib.load()
    arr = (_lib.EnsMember * 2)(_member(128, 48), _member(40, 24))
    small = _lib.EnsTuning(16384, 2048)
    big = _lib.EnsTuning(1024, 2048)
    a = L.boltx_ensemble_workspace_size(2, arr, ctypes.byref(small))
    b = L.boltx_ensemble_workspace_size(2, arr, ctypes.byref(big))
    assert 0 < a < b
    assert L.boltx_ensemble_workspace_size(2, arr, ctypes.byref(_lib.EnsTuning(16384, 64))) == a
    for bad in (_lib.EnsTuning(0, 2048), _lib.EnsTuning(16384, -1)):
        assert L.boltx_ensemble_workspace_size(2, arr, ctypes.byref(bad)) == 0
        assert b"chunk_cells" in L.boltx_last_error()
    mixed = (_lib.EnsMember * 2)(_member(128, 48, prec=0), _member(64, 24, prec=1))
    assert L.boltx_ensemble_workspace_size(2, mixed, ctypes.byref(small)) == 0
    assert b"precision" in L.boltx_last_error()
    two_d = (_lib.EnsMember * 1)(_member(64, 24))
    two_d[0].grid.n_q2 = 8
    assert L.boltx_ensemble_layout(1, two_d, None) == -1
    assert b"N_q2 = 1" in L.boltx_last_error()


def test_tuning_cache_file(tmp_path):
    """The JSON tuning cache: entries keyed by GPU model and signature,
    added without losing others, and an unreadable or malformed file (or
    entry) reads as a miss."""
    from bolt_amd import ensemble as E
    path = str(tmp_path / "sub" / "tuning.json")
    k1 = E.cache_key(((32, 8, 8, 1), (64, 6, 6, 1), 1, 1, 3), "AMD Instinct MI355X")
    k2 = E.cache_key(((32, 8, 8, 1),), "AMD Instinct MI355X")
    assert E.load_tuning(path, k1) is None
    assert E.store_tuning(path, k1, (65536, 1024))
    assert E.store_tuning(path, k2, (16384, 2048))
    assert E.load_tuning(path, k1) == (65536, 1024)
    assert E.load_tuning(path, k2) == (16384, 2048)
    assert E.load_tuning(path, E.cache_key(((32, 8, 8, 1),), "other GPU")) is None
    with open(path, "w") as fh:
        fh.write('{"%s": [1, 2, 3, 4]}' % k1.replace('"', '\\"'))
    assert E.load_tuning(path, k1) is None
    with open(path, "w") as fh:
        fh.write("not json")
    assert E.load_tuning(path, k1) is None
    assert E.store_tuning(path, k1, (4096, 1024))
    assert E.load_tuning(path, k1) == (4096, 1024)
