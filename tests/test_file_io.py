"""PETSc binary Vec files (the reference's checkpoint format,
file_io/dump.py:96-233, load.py:10-79): the wire format written byte by byte
from PETSc's published layout, round trips, and the natural ordering of the
distribution function (dof = p fastest, then q1, then q2)."""
import struct

import numpy as np
import pytest

from bolt_amd import file_io


def test_header_bytes_and_payload(tmp_path):
    v = np.array([1.0, -2.5, 3.25e-300, np.pi])
    path = str(tmp_path / "x.bin")
    file_io.write_vec(path, v, block_size=2)
    raw = open(path, "rb").read()
    # 1211214 = 0x00127B4E, big-endian, then the length 4
    assert raw[:8] == bytes([0x00, 0x12, 0x7B, 0x4E, 0, 0, 0, 4])
    assert raw[8:16] == struct.pack(">d", 1.0)
    assert raw[16:24] == bytes.fromhex("c004000000000000")  # -2.5
    assert len(raw) == 8 + 4 * 8
    assert open(path + ".info").read() == "-vecload_block_size 2\n"
    assert np.array_equal(file_io.read_vec(path), v)


def test_reader_checks_class_id_and_length(tmp_path):
    bad = tmp_path / "bad.bin"
    bad.write_bytes(struct.pack(">ii", 1211213, 1) + struct.pack(">d", 1.0))
    with pytest.raises(ValueError, match="class id"):
        file_io.read_vec(str(bad))
    short = tmp_path / "short.bin"
    short.write_bytes(struct.pack(">ii", 1211214, 3) + struct.pack(">d", 1.0))
    with pytest.raises(ValueError, match="truncated"):
        file_io.read_vec(str(short))
    ok = tmp_path / "ok.bin"
    file_io.write_vec(str(ok), np.zeros(5))
    with pytest.raises(ValueError, match="expected 6"):
        file_io.read_vec(str(ok), expected=6)


def test_natural_order_matches_af_flat_of_the_interior(tmp_path):
    """af_petsc_conversion.py:12-18: af.flat of the interior (P, S, N1, N2)
    column-major array; our dump writes the [N2][N1][P] storage in C order,
    which must be the same sequence."""
    import torch
    P, N1, N2 = 5, 3, 4
    f_ref = np.random.default_rng(0).normal(size=(P, 1, N1, N2))
    want = f_ref.reshape(-1, order="F")  # af.flat
    storage = torch.from_numpy(np.ascontiguousarray(f_ref[:, 0].transpose(2, 1, 0)))
    path = str(tmp_path / "f.bin")
    file_io.write_vec(path, file_io._natural(storage), P)
    assert np.array_equal(file_io.read_vec(path), want)


def test_restart_state_belongs_to_its_file(tmp_path):
    """The companion state of name.bin is returned only while name.bin is
    the file it was written with (its SHA-256), and only in its own format."""
    name = str(tmp_path / "f")
    file_io.write_vec(name + ".bin", np.arange(6.0), 2)
    file_io._write_state(name, {"time_elapsed": 0.125})
    st = file_io._read_state(name)
    assert st["time_elapsed"] == 0.125 and st["format"] == "bolt_amd restart state 1"
    file_io.write_vec(name + ".bin", np.arange(6.0) + 1.0, 2)
    assert file_io._read_state(name) is None
    file_io._write_state(name, {"time_elapsed": 0.5})
    assert file_io._read_state(name)["time_elapsed"] == 0.5
    open(name + ".state.json", "w").write("{not json")
    assert file_io._read_state(name) is None
    assert file_io._read_state(str(tmp_path / "missing")) is None
