"""The C-ABI library loads and exports every entry point include/attract_amd.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re
import subprocess

import numpy as np

from attract_jax_amd import _lib
from attract_jax_amd.device import LINE, SEG, lists_from_csr, pack_beads
from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "attract_amd.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(atj_\w+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    assert "atj_energrad" in names and "atj_minimize" in names and len(names) >= 10


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (atj_\w+)", out))
    for name in declared():
        assert name in exported, name
        assert hasattr(lib, name)
    assert set(_lib.SIGNATURES) == set(declared())
    assert lib.atj_version().decode().startswith("attract_amd")


def test_struct_sizes_match_c_layout():
    assert ctypes.sizeof(_lib.AtjGrid) == 4 * 3 + 4 + 12 + 12 + 4 * 5 + 4 + 8 + 8 * 5
    assert ctypes.sizeof(_lib.AtjOptimizer) == 4 * 8
    assert ctypes.sizeof(_lib.AtjJob) == 16 + 4 * 4 + 4 * 8
    assert ctypes.sizeof(_lib.AtjLigand) == 16


def test_device_list_format():
    counts = np.array([0, 3, 9, 1, 0, 8])
    starts = np.array([0, 0, 3, 12, 13, 13])
    flat = np.arange(21) * 7
    nbgrid, nblist = lists_from_csr(counts, starts, flat)
    assert np.array_equal(nbgrid[:, 0], counts)
    assert np.all(nbgrid[:, 1] % SEG == 0) and np.all(nbgrid[:, 1] >= SEG)
    assert np.all(nbgrid[counts > 0, 1] % LINE == 0)
    assert np.all(nblist[:LINE] == 0xFFFF)
    for v in range(len(counts)):
        c, s = nbgrid[v]
        assert np.array_equal(nblist[s:s + c], flat[starts[v]:starts[v] + c])
        pad_end = s + (c + SEG - 1) // SEG * SEG
        assert np.all(nblist[s + c:pad_end] == 0xFFFF)


def test_bead_packing_bit_pattern():
    b = pack_beads([[1.0, 2.0, 3.0]], [5 | (7 << 16)])
    assert b[0, :3].tolist() == [1.0, 2.0, 3.0]
    assert int(b[0, 3:].view(np.int32)[0]) == 5 | (7 << 16)
