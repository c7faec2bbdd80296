"""C ABI: the library loads without a GPU, exports every declared symbol and
rejects invalid arguments with status codes (no exceptions across the ABI)."""
import ctypes as C

import numpy as np

import cfdx_dev_amd as cx
from cfdx_dev_amd import _native


def test_every_declared_symbol_is_exported():
    lib = _native.load()
    names = _native.exported_symbols_declared()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_string():
    assert b"gfx950" in _native.load().cfdxhip_version()


def test_invalid_arguments_return_status():
    lib = _native.load()
    opt = _native.AmgOptions()
    lib.cfdxhip_amg_options_default(C.byref(opt))
    opt.method = 7
    h = C.c_void_p()
    assert lib.cfdxhip_amg_create(C.byref(opt), C.byref(h)) == 1
    lib.cfdxhip_amg_options_default(C.byref(opt))
    opt.prolongator_filter = 1.5
    assert lib.cfdxhip_amg_create(C.byref(opt), C.byref(h)) == 1


def test_host_only_handle_rejects_device_work():
    p = cx.NativeAMGPreconditioner(device=-1)
    rp = np.array([0, 2, 4], np.uint32)
    cols = np.array([0, 1, 0, 1], np.uint32)
    vals = np.array([2.0, -1.0, -1.0, 2.0])
    assert p.setup(2, rp, cols, vals), p.last_error()
    assert p.apply(np.ones(2)) is None
    assert "host-only" in p.last_error()


def test_setup_rejects_zero_diagonal():
    p = cx.NativeAMGPreconditioner(device=-1)
    rp = np.array([0, 1, 2], np.uint32)
    cols = np.array([0, 1], np.uint32)
    assert not p.setup(2, rp, cols, np.array([1.0, 0.0]))
    assert "diagonal" in p.last_error()


def test_setup_rejects_unsorted_columns():
    p = cx.NativeAMGPreconditioner(device=-1)
    rp = np.array([0, 2, 4], np.uint32)
    cols = np.array([1, 0, 0, 1], np.uint32)
    assert not p.setup(2, rp, cols, np.array([-1.0, 2.0, -1.0, 2.0]))


def test_update_values_requires_setup():
    p = cx.NativeAMGPreconditioner(device=-1)
    assert not p.update_values(np.ones(4))


def test_mesh_generator_invariants():
    m = cx.generate_mesh("box", 5, 4, 3, jitter=0.1, permute=True, seed=1)
    assert m.n_cells == 60
    assert m.n_internal_faces == 4 * 4 * 3 + 5 * 3 * 3 + 5 * 4 * 2
    assert m.n_faces == m.n_internal_faces + 2 * (4 * 3 + 5 * 3 + 5 * 4)
    nb = m.neighbour
    assert np.all(m.owner[: m.n_internal_faces] < nb[: m.n_internal_faces])
    assert np.all(nb[m.n_internal_faces:] == -1)
    # internal faces ordered by (owner, neighbour)
    key = m.owner[: m.n_internal_faces].astype(np.int64) * m.n_cells + nb[: m.n_internal_faces]
    assert np.all(np.diff(key) > 0)
    # each cell lists 6 faces, ascending
    co = m.cell_offsets
    assert np.all(np.diff(co) == 6)
    for c in range(m.n_cells):
        f = m.cell_faces[co[c]:co[c + 1]]
        assert np.all(np.diff(f) > 0)
        assert np.all((m.owner[f] == c) | (nb[f] == c))
    # permutation is a bijection of the lattice
    ijk = m.cell_ijk.reshape(-1, 3)
    assert len({tuple(r) for r in ijk}) == m.n_cells
