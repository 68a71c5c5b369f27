"""ctypes binding of libattract_amd.so (the C-ABI declared in include/attract_amd.h).

The library is the only compute path: nothing here falls back to CPU code.
Loading fails loudly when the shared library has not been built.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libattract_amd.so")

ATJ_OK = 0

ATJ_LAYOUT_PLAIN = 0
ATJ_LAYOUT_CORNER = 1

ATJ_TERM_GRID = 0x1
ATJ_TERM_NB = 0x2
ATJ_NB_PLATEAU = 0x4
ATJ_GRAD_EXACT = 0x8
ATJ_POSE_QT = 0x10
ATJ_GRAD_QT = 0x20

ATJ_NSTATS = 4
ATJ_NPHASES = 6
PHASES = ("set-up and claims", "bead phase", "segment-size scans", "pair phase", "reductions",
          "optimizer update")

c_float_p = ctypes.c_void_p  # device pointers are passed as integers


class AtjGrid(ctypes.Structure):
    _fields_ = [
        ("origin", ctypes.c_float * 3),
        ("spacing", ctypes.c_float),
        ("dim", ctypes.c_int32 * 3),
        ("dim2", ctypes.c_int32 * 3),
        ("gridextension", ctypes.c_float),
        ("plateaudis", ctypes.c_float),
        ("plateaudis2", ctypes.c_float),
        ("ntypes", ctypes.c_int32),
        ("layout", ctypes.c_int32),
        ("nshell", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("inner", ctypes.c_void_p),
        ("outer", ctypes.c_void_p),
        ("nbgrid", ctypes.c_void_p),
        ("nblist", ctypes.c_void_p),
        ("nbshell", ctypes.c_void_p),
    ]


class AtjForceField(ctypes.Structure):
    _fields_ = [
        ("rc", ctypes.c_void_p),
        ("ac", ctypes.c_void_p),
        ("ivor", ctypes.c_void_p),
        ("emin", ctypes.c_void_p),
        ("rmin2", ctypes.c_void_p),
        ("nrt", ctypes.c_int32),
        ("nlt", ctypes.c_int32),
    ]


class AtjReceptor(ctypes.Structure):
    _fields_ = [("beads", ctypes.c_void_p), ("nrec", ctypes.c_int32)]


class AtjLigand(ctypes.Structure):
    _fields_ = [("beads", ctypes.c_void_p), ("nbeads", ctypes.c_int32), ("reserved", ctypes.c_int32)]


ATJ_OPT_ADAM = 0
ATJ_OPT_ADABELIEF = 1


class AtjOptimizer(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("lr_rot", ctypes.c_float),
        ("lr_trans", ctypes.c_float),
        ("beta1", ctypes.c_float),
        ("beta2", ctypes.c_float),
        ("eps", ctypes.c_float),
        ("eps_root", ctypes.c_float),
        ("gravity", ctypes.c_float),
    ]


class AtjJob(ctypes.Structure):
    _fields_ = [
        ("ligand", AtjLigand),
        ("pose_begin", ctypes.c_int32),
        ("pose_count", ctypes.c_int32),
        ("nsteps", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("opt", AtjOptimizer),
    ]


class AtjPoseState(ctypes.Structure):
    _fields_ = [
        ("pose", ctypes.c_void_p),
        ("m", ctypes.c_void_p),
        ("v", ctypes.c_void_p),
        ("count", ctypes.c_void_p),
        ("energy", ctypes.c_void_p),
    ]


ATJ_LAUNCH_REC_GLOBAL = 0x1
ATJ_LAUNCH_CLAIM_MASK = 0xF0


def ATJ_LAUNCH_CLAIM(n):
    """Launch option: n poses per work claim (1, 2, 4 or 8)."""
    return int(n) << 4


ATJ_LAUNCH_TAIL_MASK = 0xF00
ATJ_LAUNCH_TAIL_ROUNDS_MASK = 0xF000
ATJ_LAUNCH_ADJOINT_WAVE_TABLES = 0x10000  # adjoint: per-wavefront float tables, not the fixed-point one


def ATJ_LAUNCH_TAIL(n, rounds=1):
    """Launch option: the poses of the last `rounds` rounds of claims (1..15) are claimed n at a
    time (1, 2 or 4, below the claim size)."""
    return (int(n) << 8) | (int(rounds) << 12)


class AtjLaunch(ctypes.Structure):
    _fields_ = [("waves_per_block", ctypes.c_int32), ("blocks", ctypes.c_int32), ("options", ctypes.c_uint32)]


# (name, argtypes) of every exported entry point; tests check the exports against this table
_P = ctypes.c_void_p
_I = ctypes.c_int32
_U = ctypes.c_uint32
_F = ctypes.c_float
SIGNATURES = {
    "atj_last_error": ([], ctypes.c_char_p),
    "atj_version": ([], ctypes.c_char_p),
    "atj_energrad": ([ctypes.POINTER(AtjGrid), ctypes.POINTER(AtjForceField), ctypes.POINTER(AtjReceptor),
                      _P, _I, _P, _P, _P, _I, _P, _P, _U, ctypes.POINTER(AtjLaunch), _P, _P, _P], ctypes.c_int),
    "atj_point_terms": ([ctypes.POINTER(AtjGrid), ctypes.POINTER(AtjForceField), ctypes.POINTER(AtjReceptor),
                         _P, _P, _I, _P, _P, _P, _P, _U, _P], ctypes.c_int),
    "atj_bead_terms": ([ctypes.POINTER(AtjGrid), ctypes.POINTER(AtjForceField), ctypes.POINTER(AtjReceptor),
                        ctypes.POINTER(AtjLigand), _P, _I, _P, _P, _P, _P, _U, _P], ctypes.c_int),
    "atj_minimize": ([ctypes.POINTER(AtjGrid), ctypes.POINTER(AtjForceField), ctypes.POINTER(AtjReceptor),
                      _P, _I, _P, _P, _I, ctypes.POINTER(AtjPoseState), _I, _U, ctypes.POINTER(AtjLaunch),
                      _P, _P, _P], ctypes.c_int),
    "atj_minimize_noalloc": ([ctypes.POINTER(AtjGrid), ctypes.POINTER(AtjForceField), ctypes.POINTER(AtjReceptor),
                              _P, _I, _P, _P, _I, ctypes.POINTER(AtjPoseState), _I, _U, ctypes.POINTER(AtjLaunch),
                              _P, _P], ctypes.c_int),
    "atj_ligand_rmsd": ([_P, _I, _P, _P, _I, _P, _P], ctypes.c_int),
    "atj_optimizer_step": ([_P, _P, _I, ctypes.POINTER(AtjPoseState), ctypes.POINTER(AtjOptimizer), _P],
                           ctypes.c_int),
    "atj_energrad_adjoint": ([ctypes.POINTER(AtjGrid), ctypes.POINTER(AtjForceField),
                              ctypes.POINTER(AtjReceptor), _P, _I, _P, _P, _I, _P, _P, _P, _P, _U,
                              ctypes.POINTER(AtjLaunch), _P, _P], ctypes.c_int),
    "atj_reduce_partials": ([_P, _I, _I, _P, _P], ctypes.c_int),
    "atj_grid_build_potentials": ([ctypes.POINTER(AtjGrid), ctypes.POINTER(AtjForceField),
                                   ctypes.POINTER(AtjReceptor), _P, _F, _P, _P, _P], ctypes.c_int),
    "atj_grid_count_neighbours": ([ctypes.POINTER(AtjGrid), ctypes.POINTER(AtjReceptor), _F, _P, _P],
                                  ctypes.c_int),
    "atj_grid_fill_neighbours": ([ctypes.POINTER(AtjGrid), ctypes.POINTER(AtjReceptor), _F, _P, _P, _P, _P, _P],
                                 ctypes.c_int),
    "atj_grid_shell_lists": ([ctypes.POINTER(AtjGrid), ctypes.POINTER(AtjReceptor), _I, _P, _P, _P], ctypes.c_int),
    "atj_grid_pack_corners": ([_P, _I, ctypes.POINTER(ctypes.c_int32 * 3), _P, _P], ctypes.c_int),
}

_lib = None


class AttractError(RuntimeError):
    pass


def load():
    """Load libattract_amd.so once; raise if it is missing (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    path = os.environ.get("ATTRACT_AMD_LIB", LIB_PATH)  # another build of the same library
    if not os.path.exists(path):
        raise AttractError(
            f"{path} is not built: run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    lib = ctypes.CDLL(path)
    for name, (argtypes, restype) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = restype
    _lib = lib
    return lib


def check(status):
    if status != ATJ_OK:
        msg = _lib.atj_last_error().decode() if _lib is not None else "?"
        raise AttractError(f"libattract_amd error {status}: {msg}")
