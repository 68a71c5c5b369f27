"""ctypes wrapper of the FP32 CPU twin (oracle/twin.c).  TEST INFRASTRUCTURE ONLY."""
import ctypes
import os
import subprocess

import numpy as np

from . import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libattract_twin.so")

T_GRID, T_NB, T_PLATEAU, T_EXACT, T_QT = 0x1, 0x2, 0x4, 0x8, 0x10


class TwinGrid(ctypes.Structure):
    _fields_ = [("origin", ctypes.c_float * 3), ("spacing", ctypes.c_float), ("dim", ctypes.c_int * 3),
                ("dim2", ctypes.c_int * 3), ("gridextension", ctypes.c_float), ("plateaudis2", ctypes.c_float),
                ("ntypes", ctypes.c_int), ("inner", ctypes.c_void_p), ("outer", ctypes.c_void_p),
                ("counts", ctypes.c_void_p), ("starts", ctypes.c_void_p), ("lists", ctypes.c_void_p)]


class TwinFF(ctypes.Structure):
    _fields_ = [("rc", ctypes.c_void_p), ("ac", ctypes.c_void_p), ("ivor", ctypes.c_void_p),
                ("emin", ctypes.c_void_p), ("rmin2", ctypes.c_void_p), ("nrt", ctypes.c_int), ("nlt", ctypes.c_int)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def pack4(xyz, tag):
    out = np.empty((len(xyz), 4), np.float32)
    out[:, :3] = np.asarray(xyz, np.float32)
    out[:, 3] = np.asarray(tag, np.int32).view(np.float32)
    return out


class Twin:
    """A reference Grid + pair table + receptor + ligand held for twin calls."""

    def __init__(self, grid, ff, coor_rec, rec_fftypes, coor_lig, gtypes, lig_fftypes):
        self.keep = []
        counts, starts, lists = oracle.csr_lists(grid)
        k = lambda a: self.keep.append(a) or a
        g = TwinGrid()
        g.origin[:] = [float(np.float32(v)) for v in grid.origin]
        g.spacing = float(np.float32(grid.gridspacing))
        g.dim[:] = [int(v) for v in grid.dim]
        g.dim2[:] = [int(v) for v in grid.dim2]
        g.gridextension = float(grid.gridextension)
        g.plateaudis2 = float(np.float32(float(grid.plateaudis) ** 2))
        inner = getattr(grid, "inner_potential_grid", None)
        outer = getattr(grid, "outer_potential_grid", None)
        g.inner = _p(k(np.ascontiguousarray(inner, np.float32))) if inner is not None else None
        g.outer = _p(k(np.ascontiguousarray(outer, np.float32))) if outer is not None else None
        g.ntypes = int(inner.shape[0]) if inner is not None else 0
        g.counts = _p(k(np.ascontiguousarray(np.maximum(counts, 0), np.int32)))
        g.starts = _p(k(np.ascontiguousarray(np.maximum(starts, 0), np.int64)))
        lists = np.asarray(lists, np.int64)
        g.lists = _p(k(np.ascontiguousarray(np.where(lists >= 65535, 0, lists), np.int32)))
        self.grid = g
        f = TwinFF()
        arrs = [k(np.ascontiguousarray(np.asarray(a, np.float64).astype(np.float32))) for a in ff]
        f.rc, f.ac, f.ivor, f.emin, f.rmin2 = [_p(a) for a in arrs]
        f.nrt, f.nlt = arrs[0].shape
        self.ff = f
        self.rec4 = k(pack4(coor_rec, rec_fftypes))
        self.lig4 = k(pack4(coor_lig, np.asarray(gtypes, np.int64) | (np.asarray(lig_fftypes, np.int64) << 16)))
        self.nb = len(self.lig4)
        self.has_pot = inner is not None

    def default_flags(self):
        return (T_GRID | T_NB | T_PLATEAU) if self.has_pot else T_NB

    def energrad(self, poses, flags=None, qt=False, nthreads=0):
        lib = load()
        P = np.ascontiguousarray(poses, np.float32).reshape(len(poses), -1)
        N = len(P)
        fl = (self.default_flags() if flags is None else flags) | (T_QT if qt else 0)
        E = np.zeros(N, np.float32)
        G = np.zeros((N, 16), np.float32)
        lib.twin_energrad(ctypes.byref(self.grid), ctypes.byref(self.ff), _p(self.rec4), _p(self.lig4),
                          ctypes.c_int(self.nb), _p(P), ctypes.c_int(N), ctypes.c_uint(fl), _p(E), _p(G),
                          ctypes.c_int(nthreads))
        return E, G.reshape(N, 4, 4)

    def bead_terms(self, poses, flags=None, qt=False, nthreads=0):
        lib = load()
        P = np.ascontiguousarray(poses, np.float32).reshape(len(poses), -1)
        N, nb = len(P), self.nb
        fl = (self.default_flags() if flags is None else flags) | (T_QT if qt else 0)
        out = dict(grid4=np.zeros((N, nb, 4), np.float32), voxel=np.zeros((N, nb), np.int32),
                   nbene=np.zeros((N, nb), np.float32), nbgrad=np.zeros((N, nb, 3), np.float32),
                   npairs=np.zeros((N, nb), np.int32), xp=np.zeros((N, nb, 3), np.float32))
        lib.twin_bead_terms(ctypes.byref(self.grid), ctypes.byref(self.ff), _p(self.rec4), _p(self.lig4),
                            ctypes.c_int(nb), _p(P), ctypes.c_int(N), ctypes.c_uint(fl), _p(out["grid4"]),
                            _p(out["voxel"]), _p(out["nbene"]), _p(out["nbgrad"]), _p(out["npairs"]),
                            _p(out["xp"]), ctypes.c_int(nthreads))
        return out

    def minimize(self, pose7, nsteps, lr_rot, lr_trans, beta1=0.9, beta2=0.999, eps=1e-8, flags=None, nthreads=0,
                 kind=0, eps_root=0.0, gravity=0.0, bias_pow=False):
        """kind 0 = optax.adam, 1 = optax.adabelief; gravity = restraint weight; bias_pow: bias
        corrections by powf instead of the recurrence the GPU minimizer uses."""
        lib = load()
        P = np.array(pose7, np.float32, order="C").reshape(-1, 7)
        N = len(P)
        fl = self.default_flags() if flags is None else flags
        E = np.zeros(N, np.float32)
        lib.twin_minimize(ctypes.byref(self.grid), ctypes.byref(self.ff), _p(self.rec4), _p(self.lig4),
                          ctypes.c_int(self.nb), _p(P), ctypes.c_int(N), ctypes.c_int(int(nsteps)),
                          ctypes.c_float(lr_rot), ctypes.c_float(lr_trans), ctypes.c_float(beta1),
                          ctypes.c_float(beta2), ctypes.c_float(eps), ctypes.c_int(int(kind)),
                          ctypes.c_float(eps_root), ctypes.c_float(gravity), ctypes.c_uint(fl), _p(E),
                          ctypes.c_int(nthreads), ctypes.c_int(int(bool(bias_pow))))
        return P, E
