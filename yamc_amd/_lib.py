"""ctypes binding of libyamc_ewald.so (C ABI in include/yamc_ewald.h).

The library is loaded from the package directory.  There is no fallback: if
the shared object is missing or cannot be loaded, YamcLibraryError is raised.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("YAMC_LIB", os.path.join(HERE, "libyamc_ewald.so"))

YAMC_FP64 = 0
YAMC_FP32C = 1
MAX_MOL_ATOM = 4
MAX_KMAX = 40

ERRORS = {1: "ERROR_IO", 2: "ERROR_LIMITS", 3: "ERROR_PARAMETER", 4: "ERROR_INITIALIZATION",
          5: "ERROR_WRONG_FUNCTION", 6: "ERROR_DEVICE"}

EXPORTS = [
    "yamc_ctx_create", "yamc_ctx_destroy", "yamc_ctx_info", "yamc_set_positions", "yamc_get_positions",
    "yamc_grid", "yamc_kspace_init", "yamc_kspace_get_beta", "yamc_kspace_get_sums", "yamc_kspace_delta",
    "yamc_kspace_forces", "yamc_rspace_full", "yamc_rspace_full_noalloc", "yamc_rspace_partial", "yamc_tails", "yamc_external_mu",
    "yamc_mc_init", "yamc_mc_energies", "yamc_mc_real_arrays", "yamc_mc_set_rng", "yamc_mc_set_rng_list",
    "yamc_mc_get_rng",
    "yamc_mc_run", "yamc_mc_run_const", "yamc_resync", "yamc_mc_volume_move", "yamc_mc_run_npt", "yamc_box_length",
    "yamc_mc_set_frequencies", "yamc_mc_xchange", "yamc_set_box", "yamc_mc_commit_move", "yamc_batch_run", "yamc_batch_create", "yamc_batch_launch",
    "yamc_batch_destroy", "yamc_forces", "yamc_rdf",
    "yamc_set_block_size", "yamc_set_chain_groups", "yamc_set_cell_div", "yamc_last_kernel_ms", "yamc_phase_times", "yamc_last_error",
    # per-procedure surface of the Fortran drop-in modules
    "yamc_ws_create", "yamc_ws_destroy", "yamc_ws_grid", "yamc_op_beta", "yamc_op_beta_lj", "yamc_op_sincos_xyz",
    "yamc_op_sincos_kr", "yamc_op_sincos_kr_atom", "yamc_op_sums", "yamc_op_rho_squared", "yamc_op_energy_coulomb",
    "yamc_op_energy_lj", "yamc_op_du_coulomb", "yamc_op_du_lj", "yamc_op_force_coulomb", "yamc_op_force_lj",
    "yamc_op_rspace", "yamc_op_pair", "yamc_op_tails",
]


class YamcLibraryError(RuntimeError):
    pass


class YamcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class SystemDesc(C.Structure):
    _fields_ = [
        ("natom", C.c_int32), ("nmol", C.c_int32), ("ntype", C.c_int32), ("kmax", C.c_int32),
        ("precision", C.c_int32), ("pad", C.c_int32),
        ("box_length", C.c_double), ("alpha", C.c_double), ("rmax2", C.c_double), ("xclb", C.c_double),
        ("kext", C.c_double), ("dr", C.c_double), ("d_angle", C.c_double), ("lambda_f", C.c_double),
        ("lambda_c", C.c_double),
        ("charge", C.POINTER(C.c_double)), ("hard_core_angstr", C.POINTER(C.c_double)),
        ("mass", C.POINTER(C.c_double)),
        ("molnum", C.POINTER(C.c_int32)), ("lj_type", C.POINTER(C.c_int32)),
        ("mol_first", C.POINTER(C.c_int32)), ("mol_last", C.POINTER(C.c_int32)),
        ("lj6", C.POINTER(C.c_double)), ("lj12", C.POINTER(C.c_double)), ("four_eps", C.POINTER(C.c_double)),
        ("sigma6", C.POINTER(C.c_double)), ("tail_counts", C.POINTER(C.c_int32)),
    ]


class MoveRec(C.Structure):
    _fields_ = [
        ("imol", C.c_int32), ("accepted", C.c_int32), ("overlap", C.c_int32), ("imol2", C.c_int32),
        ("du_real", C.c_double), ("du_kc", C.c_double), ("du_klj", C.c_double), ("du_ext", C.c_double),
        ("du_total", C.c_double),
        ("f_old", C.c_double * 3), ("t_old", C.c_double * 3), ("f_new", C.c_double * 3), ("t_new", C.c_double * 3),
        ("delta", C.c_double * 3), ("angle", C.c_double),
    ]


MOVE_DTYPE = np.dtype([
    ("imol", "<i4"), ("accepted", "<i4"), ("overlap", "<i4"), ("imol2", "<i4"),
    ("du_real", "<f8"), ("du_kc", "<f8"), ("du_klj", "<f8"), ("du_ext", "<f8"), ("du_total", "<f8"),
    ("f_old", "<f8", (3,)), ("t_old", "<f8", (3,)), ("f_new", "<f8", (3,)), ("t_new", "<f8", (3,)),
    ("delta", "<f8", (3,)), ("angle", "<f8"),
])
assert MOVE_DTYPE.itemsize == C.sizeof(MoveRec)

_lib: Optional[C.CDLL] = None


def load(path: str = LIB_PATH) -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise YamcLibraryError(f"{path} is not built: run __graft_entry__.build() (the HIP library has no fallback)")
    try:
        lib = C.CDLL(path)
    except OSError as e:  # pragma: no cover - depends on the ROCm runtime
        raise YamcLibraryError(f"cannot load {path}: {e}") from e
    for name in EXPORTS:
        if not hasattr(lib, name):
            raise YamcLibraryError(f"{path} does not export {name}")
    lib.yamc_last_error.restype = C.c_char_p
    for name in EXPORTS:
        if name != "yamc_last_error":
            getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != 0:
        msg = _lib.yamc_last_error().decode() if _lib is not None else ""
        raise YamcError(rc, msg)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int32))
