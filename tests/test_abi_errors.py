"""Argument validation of the C-ABI: invalid grids, pair tables, sizes and NULL buffers are
refused with ATJ_ERR_INVALID / ATJ_ERR_UNSUPPORTED and a message from atj_last_error(),
before any device call (these run without a GPU; the pointers passed are host addresses
that the library must not touch when it refuses a call)."""
import ctypes

import numpy as np
import pytest

from attract_jax_amd import _lib

INVALID, UNSUPPORTED = 1, 3
_keep = []  # host buffers whose addresses stand in for device pointers


def _addr(n=64, dtype=np.float32):
    a = np.zeros(n, dtype)
    _keep.append(a)
    return a.ctypes.data


def _grid(**kw):
    g = _lib.AtjGrid()
    g.origin[:] = [0.0, 0.0, 0.0]
    g.spacing = 0.9
    g.dim[:] = [8, 8, 8]
    g.dim2[:] = [8, 8, 8]
    g.gridextension = 4.0
    g.plateaudis = 6.0
    g.plateaudis2 = 36.0
    g.ntypes = 1
    g.layout = _lib.ATJ_LAYOUT_CORNER
    g.inner = _addr()
    g.outer = 0
    g.nbgrid = _addr(dtype=np.int32)
    g.nblist = _addr(dtype=np.uint16)
    for k, v in kw.items():
        if k in ("dim", "dim2"):
            getattr(g, k)[:] = v
        else:
            setattr(g, k, v)
    return g


def _ff(nrt=2, nlt=2):
    return _lib.AtjForceField(_addr(), _addr(), _addr(), _addr(), _addr(), nrt, nlt)


def _rec(n=10):
    return _lib.AtjReceptor(_addr(4 * n), n)


def _energrad(lib, grid, ff=None, rec=None, N=4, poses=True, energy=True, flags=None, launch=None):
    ff = ff or _ff()
    rec = rec or _rec()
    flags = _lib.ATJ_TERM_GRID | _lib.ATJ_TERM_NB if flags is None else flags
    launch = launch or _lib.AtjLaunch(16, 0, 0)
    return lib.atj_energrad(ctypes.byref(grid) if grid is not None else None, ctypes.byref(ff), ctypes.byref(rec),
                            _addr(), 1, None, None, _addr() if poses else None, N, _addr() if energy else None,
                            None, flags, ctypes.byref(launch), _addr(dtype=np.int32), None, None)


@pytest.fixture(scope="module")
def lib():
    return _lib.load()


def _msg(lib):
    return lib.atj_last_error().decode()


def test_null_grid(lib):
    assert _energrad(lib, None) == INVALID and "grid" in _msg(lib)


def test_grid_dims_spacing_layout(lib):
    assert _energrad(lib, _grid(dim=[1, 8, 8])) == INVALID and "dim" in _msg(lib)
    assert _energrad(lib, _grid(spacing=0.0)) == INVALID and "spacing" in _msg(lib)
    assert _energrad(lib, _grid(layout=7)) == INVALID and "layout" in _msg(lib)


def test_neighbour_lists_required_for_the_pair_term(lib):
    assert _energrad(lib, _grid(nblist=0)) == INVALID and "neighbour lists" in _msg(lib)


def test_shell_count_range_and_entries(lib):
    assert _energrad(lib, _grid(nshell=7)) == INVALID and "nshell" in _msg(lib)
    assert _energrad(lib, _grid(nshell=-1)) == INVALID and "nshell" in _msg(lib)
    assert _energrad(lib, _grid(nshell=4, nbshell=0)) == INVALID and "nbshell" in _msg(lib)


def test_potential_grid_count(lib):
    assert _energrad(lib, _grid(ntypes=0)) == INVALID and "ntypes" in _msg(lib)
    assert _energrad(lib, _grid(ntypes=70000)) == INVALID and "ntypes" in _msg(lib)


def test_grid_builder_needs_a_pair_table(lib):
    st = lib.atj_grid_build_potentials(ctypes.byref(_grid()), ctypes.byref(_ff(0, 2)), ctypes.byref(_rec()),
                                       _addr(dtype=np.int32), 50.0, _addr(), None, None)
    assert st == INVALID and "pair table" in _msg(lib)


def test_grid_term_needs_potentials(lib):
    assert _energrad(lib, _grid(inner=0), flags=_lib.ATJ_TERM_GRID) == INVALID and "potentials" in _msg(lib)


def test_pair_table_and_receptor_limits(lib):
    assert _energrad(lib, _grid(), ff=_ff(0, 2)) == INVALID and "pair table" in _msg(lib)
    assert _energrad(lib, _grid(), ff=_ff(65, 64)) == UNSUPPORTED and "4096" in _msg(lib)
    # 65536 x 65536 is 0 in 32-bit arithmetic: the size check must not wrap
    assert _energrad(lib, _grid(), ff=_ff(65536, 65536)) == UNSUPPORTED and "4096" in _msg(lib)
    assert _energrad(lib, _grid(), rec=_lib.AtjReceptor(_addr(), 70000)) == UNSUPPORTED and "65536" in _msg(lib)
    assert _energrad(lib, _grid(), rec=_lib.AtjReceptor(None, 10)) == INVALID and "receptor" in _msg(lib)


def test_pose_buffers_and_counts(lib):
    assert _energrad(lib, _grid(), N=-1) == INVALID
    assert _energrad(lib, _grid(), poses=False) == INVALID and "NULL" in _msg(lib)
    assert _energrad(lib, _grid(), energy=False) == INVALID and "NULL" in _msg(lib)
    assert _energrad(lib, _grid(), N=0) == 0  # nothing to do: no device call


def test_shell_lists_arguments(lib):
    rec = _rec()
    out16, out32 = _addr(dtype=np.uint16), _addr(dtype=np.int32)
    for n in (0, 7):
        st = lib.atj_grid_shell_lists(ctypes.byref(_grid()), ctypes.byref(rec), n, out16, out32, None)
        assert st == INVALID and "nshell" in _msg(lib)
    st = lib.atj_grid_shell_lists(ctypes.byref(_grid(nshell=4, nbshell=_addr())), ctypes.byref(rec), 4, out16, out32,
                                  None)
    assert st == INVALID and "nshell 0" in _msg(lib)
    st = lib.atj_grid_shell_lists(ctypes.byref(_grid()), ctypes.byref(rec), 4, None, out32, None)
    assert st == INVALID and "NULL" in _msg(lib)


def test_minimize_counts(lib):
    state = _lib.AtjPoseState(_addr(), _addr(), _addr(), _addr(dtype=np.int32), _addr())
    launch = _lib.AtjLaunch(16, 0, 0)
    st = lib.atj_minimize(ctypes.byref(_grid()), ctypes.byref(_ff()), ctypes.byref(_rec()), _addr(), 1, None, None, 4,
                          ctypes.byref(state), -1, _lib.ATJ_TERM_GRID | _lib.ATJ_TERM_NB, ctypes.byref(launch),
                          _addr(dtype=np.int32), None, None)
    assert st == INVALID and "nsteps" in _msg(lib)
    st = lib.atj_minimize(ctypes.byref(_grid()), ctypes.byref(_ff()), ctypes.byref(_rec()), _addr(), 1, None, None, 4,
                          ctypes.byref(state), 0, _lib.ATJ_TERM_GRID | _lib.ATJ_TERM_NB, ctypes.byref(launch),
                          _addr(dtype=np.int32), None, None)
    assert st == 0  # zero steps: nothing to do


def test_launch_options(lib):
    """Waves per block, unknown option bits and claim sizes are refused before any device call."""
    L = lambda w, o: _lib.AtjLaunch(w, 0, o)
    assert _energrad(lib, _grid(), launch=L(17, 0)) == INVALID and "waves_per_block" in _msg(lib)
    assert _energrad(lib, _grid(), launch=L(16, 1 << 20)) == INVALID and "unknown launch option" in _msg(lib)
    for n in (3, 5, 6, 15):
        assert _energrad(lib, _grid(), launch=L(16, _lib.ATJ_LAUNCH_CLAIM(n))) == INVALID, n
        assert "per claim" in _msg(lib)
    for n in (3, 8):
        assert _energrad(lib, _grid(), launch=L(16, _lib.ATJ_LAUNCH_CLAIM(8) | _lib.ATJ_LAUNCH_TAIL(n))) == INVALID, n
        assert "tail claim" in _msg(lib)
