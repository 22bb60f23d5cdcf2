"""Checkpoints in the reference's wire format: PETSc binary Vec files.

bolt/lib/nonlinear/file_io/dump.py:96-233 and load.py:10-79 write and read
the DMDA global vectors of f, of the moments and of the Yee-grid fields with
``PETSc.Viewer().createBinary(name + '.bin')``.  That file is:

    int32  VEC_FILE_CLASSID = 1211214   (big-endian)
    int32  N                            (big-endian, global length)
    N x float64                         (big-endian IEEE)

in the DMDA's natural ordering: the dof index fastest, then q1, then q2,
where the dof of f is p + P*s (af.flat of the interior (P, S, N_q1, N_q2)
array, af_petsc_conversion.py:3-20).  PETSc also writes ``name + '.bin.info'``
holding ``-vecload_block_size <dof>`` for vectors with more than one dof per
point; it is written here too and ignored on load.

Only files whose scalars are float64 and whose integers are 32-bit (PETSc's
defaults, which the reference builds with) are produced and accepted.
"""
import os
import struct

import numpy as np
import torch

VEC_FILE_CLASSID = 1211214
F64 = torch.float64


def write_vec(file_name, values, block_size=1):
    """Write a 1-D float64 array as a PETSc binary Vec (``file_name`` is the
    full path, ``.bin`` included)."""
    a = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(-1))
    with open(file_name, 'wb') as fh:
        fh.write(struct.pack('>ii', VEC_FILE_CLASSID, a.size))
        fh.write(a.astype('>f8').tobytes())
    if block_size > 1:
        with open(file_name + '.info', 'w') as fh:
            fh.write('-vecload_block_size %d\n' % block_size)


def read_vec(file_name, expected=None):
    """Read a PETSc binary Vec written by write_vec or by PETSc."""
    with open(file_name, 'rb') as fh:
        head = fh.read(8)
        if len(head) != 8:
            raise ValueError('%s: not a PETSc binary Vec (short header)' % file_name)
        cid, n = struct.unpack('>ii', head)
        if cid != VEC_FILE_CLASSID:
            raise ValueError('%s: class id %d is not a Vec (%d)' % (file_name, cid, VEC_FILE_CLASSID))
        if expected is not None and n != expected:
            raise ValueError('%s holds %d values, expected %d' % (file_name, n, expected))
        data = fh.read(8 * n)
        if len(data) != 8 * n:
            raise ValueError('%s: truncated (%d of %d values)' % (file_name, len(data) // 8, n))
    return np.frombuffer(data, dtype='>f8').astype(np.float64)


def _natural(x_qqd):
    """[N_q2][N_q1][dof] tensor -> flat natural-order numpy array."""
    return x_qqd.detach().to(F64).contiguous().cpu().numpy().reshape(-1)


def moment_names(solver):
    """nonlinear_solver.py:284-288: the public attributes of the moments
    module in dir() order, without integral_over_v."""
    names = [a for a in dir(solver.physical_system.moments) if not a.startswith('_')]
    if 'integral_over_v' in names:
        names.remove('integral_over_v')
    return names


def dump_distribution_function(solver, file_name):
    """dump.py:121-168: f's interior in natural order (dof = p + P s fastest,
    then q1, then q2), to file_name + '.bin'."""
    s = solver
    parts = []
    for sp in s._spc:
        f = sp.f if s._layout == 0 else sp.f.permute(0, 2, 1)  # -> [N_q2][N_q1][P]
        parts.append(f)
    f = torch.stack(parts, dim=2)  # [N_q2][N_q1][S][P]
    write_vec(file_name + '.bin', _natural(f), s._P * s.N_species)


def load_distribution_function(solver, file_name):
    """load.py:10-43: f from file_name + '.bin' (ghost zones refilled by the
    boundary conditions before the next RHS)."""
    s = solver
    S = s.N_species
    v = read_vec(file_name + '.bin', s._Q * s._P * S)
    x = torch.from_numpy(v).to(s.device).reshape(s.N_q2, s.N_q1, S, s._P)
    for k in s._species_loop():
        y = x[:, :, k]
        if s._layout != 0:
            y = y.permute(0, 2, 1)
        s._f = s._store(y.to(s._dtype))
        s._update_moments()


def dump_moments(solver, file_name):
    """dump.py:96-119: every moment of the moments module, joined along the
    species axis (attribute slowest), in natural order."""
    s = solver
    cols = []
    for name in moment_names(s):
        m = s.compute_moments(name)  # (1, S, Q1g, Q2g)
        cols.append(s._q_interior(m)[0])  # (S, N_q1, N_q2)
    x = torch.cat(cols, dim=0)  # (n_moments * S, N_q1, N_q2)
    write_vec(file_name + '.bin', _natural(x.permute(2, 1, 0)), x.shape[0])


def dump_EM_fields(solver, file_name):
    """dump.py:170-233: the Yee-grid fields E1, E2, E3, B1, B2, B3."""
    fs = solver.fields_solver
    if fs is None:
        raise AttributeError('fields are not enabled: there is no fields solver to dump')
    write_vec(file_name + '.bin', _natural(fs._yee.t()), 6)


def load_EM_fields(solver, file_name):
    """load.py:45-79: the Yee-grid fields, then their cell-centred values."""
    fs = solver.fields_solver
    if fs is None:
        raise AttributeError('fields are not enabled: there is no fields solver to load into')
    v = read_vec(file_name + '.bin', 6 * solver._Q)
    fs._yee.copy_(torch.from_numpy(v).to(solver.device).reshape(solver._Q, 6).t())
    fs._yee_to_cc(fs._cc)


def dump_aux_arrays(solver, arrays, name, file_name):
    """dump.py:9-39 writes HDF5 through PETSc; HDF5 is not available here."""
    raise NotImplementedError('dump_aux_arrays writes HDF5 (PETSc HDF5 viewer); no HDF5 '
                              'library is available to this package')


def exists(file_name):
    return os.path.exists(file_name + '.bin')
