"""ctypes bindings of libpantea_hip.so (the C ABI declared in include/pantea_hip.h).

The library is required: importing this module succeeds on machines without
a GPU (so that metadata and host logic can be tested), but every entry point
that launches work checks for the library and for a HIP device first and
raises if either is missing.  There is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path
from typing import Optional

PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("PANTEA_HIP_LIB", PKG / "libpantea_hip.so"))

PT_OK = 0
PT_ERR_ARG = -1
PT_ERR_CAPACITY = -2
PT_ERR_HIP = -3
PT_ERR_UNSUPPORTED = -4

PT_F64 = 0
PT_F32 = 1
PT_ACSF_KIND_SHIFT = 8
PT_ACSF_GENERIC, PT_ACSF_TANH_P4, PT_ACSF_TANH_Z1N2, PT_ACSF_TANH_Z14 = 0, 1, 2, 3

PT_MLP_F64 = 0
PT_MLP_BF16 = 1
PT_MLP_BF16X3 = 2

CUTOFF_CODES = {"hard": 0, "cos": 1, "tanhu": 2, "tanh": 3, "exp": 4, "poly1": 5, "poly2": 6}
ACTIVATION_CODES = {
    "identity": 0, "tanh": 1, "logistic": 2, "softplus": 3, "relu": 4,
    "gaussian": 5, "cos": 6, "exp": 7, "harmonic": 8,
}

PT_RG_MAX = 8
PT_AG_ETA = 4
PT_AG_ANG = 4
PT_MAX_GROUPS = 24

_c_int = ctypes.c_int
_c_dbl = ctypes.c_double
_vp = ctypes.c_void_p
_c_ll = ctypes.c_longlong
_c_size = ctypes.c_size_t


class RadialGroup(ctypes.Structure):
    _fields_ = [
        ("elem", _c_int), ("cut", _c_int), ("n", _c_int), ("pad_", _c_int),
        ("rc", _c_dbl),
        ("eta", _c_dbl * PT_RG_MAX), ("rs", _c_dbl * PT_RG_MAX),
        ("col", _c_int * PT_RG_MAX),
    ]


class AngularGroup(ctypes.Structure):
    _fields_ = [
        ("ej", _c_int), ("ek", _c_int), ("cut", _c_int), ("kind", _c_int),
        ("n_eta", _c_int), ("n_ang", _c_int),
        ("rc", _c_dbl), ("half", _c_dbl),
        ("eta", _c_dbl * PT_AG_ETA),
        ("zeta", _c_dbl * PT_AG_ANG), ("lam", _c_dbl * PT_AG_ANG), ("pre", _c_dbl * PT_AG_ANG),
        ("izeta", _c_int * PT_AG_ANG),
        ("col", _c_int * (PT_AG_ETA * PT_AG_ANG)),
    ]


class SFTable(ctypes.Structure):
    _fields_ = [
        ("n_sf", _c_int), ("n_rg", _c_int), ("n_ag", _c_int), ("pad_", _c_int),
        ("rc_max", _c_dbl), ("tanh_pre", _c_dbl),
        ("rg", RadialGroup * PT_MAX_GROUPS),
        ("ag", AngularGroup * PT_MAX_GROUPS),
    ]


class MlpDesc(ctypes.Structure):
    _fields_ = [
        ("n_layers", _c_int), ("in_dim", _c_int),
        ("width", _c_int * 4), ("act", _c_int * 4),
        ("scale_add", _c_dbl),
    ]


class PermuteArray(ctypes.Structure):
    _fields_ = [("buf", _vp), ("width", _c_int), ("word_bytes", _c_int)]


class HipLibraryError(RuntimeError):
    pass


_SIGS = {
    "pt_last_error": (ctypes.c_char_p, []),
    "pt_version": (_c_int, []),
    "pt_nbr_workspace_bytes": (_c_size, [_c_int, _c_int]),
    "pt_nbr_build": (_c_int, [_c_int, _vp, _c_int, _vp, _c_int, _vp, _c_dbl, _vp, _vp, _vp, _vp, _vp,
                              _c_ll, ctypes.POINTER(_c_ll), ctypes.POINTER(_c_int), _vp, _c_size, _vp]),
    "pt_nbr_build_async": (_c_int, [_c_int, _vp, _c_int, _vp, _c_int, _vp, _c_dbl, _vp, _vp, _vp, _vp, _vp,
                                    _c_ll, _vp, _vp, _c_int, _vp, _c_size, _vp]),
    "pt_nbr_mark_rows": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp]),
    "pt_skin_check": (_c_int, [_c_int, _vp, _vp, _vp, _c_dbl, _vp, _vp, _vp]),
    "pt_copy_gated": (_c_int, [_c_ll, _c_int, _vp, _vp, _vp, _vp]),
    "pt_permute_rows_gated": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp]),
    "pt_sf_table_kind": (_c_int, [ctypes.POINTER(SFTable)]),
    "pt_acsf_forward": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp, _vp,
                                 _vp, _c_int, _vp]),
    "pt_acsf_adjoint": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp, _vp,
                                 _vp, _c_int, _vp, _vp, _vp]),
    "pt_acsf_tangent": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp, _vp,
                                 _vp, _c_int, _vp, _c_int, _vp]),
    "pt_acsf_adjoint_transposed": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp, _vp,
                                            _vp, _c_int, _vp, _vp, _vp]),
    "pt_force_gather": (_c_int, [_c_int, _vp, _vp, _vp, _c_int, _vp, _vp]),
    "pt_force_gather_transposed": (_c_int, [_c_int, _vp, _vp, _vp, _c_int, _vp, _vp]),
    "pt_mlp_packed_bytes": (_c_size, [_vp]),
    "pt_mlp_pack": (_c_int, [_vp, _vp, _vp, _vp]),
    "pt_mlp_energy_grad": (_c_int, [_c_int, _c_int, _vp, _c_int, _c_int, _vp, _vp, _vp, _vp,
                                    _vp, _vp]),
    "pt_scaler_workspace_bytes": (_c_size, [_c_int, _c_int]),
    "pt_scaler_partial_fit": (_c_int, [_c_int, _c_int, _vp, _c_int, _c_int, _c_dbl, _vp, _vp, _c_size, _vp]),
    "pt_scaler_extrapolation": (_c_int, [_c_int, _c_int, _vp, _c_int, _c_int, _vp, _c_dbl, _vp, _vp]),
    "pt_reduce_workspace_bytes": (_c_size, [_c_int]),
    "pt_reduce_sum_f64": (_c_int, [_c_int, _vp, _vp, _vp, _c_size, _vp]),
    "pt_reduce_sum_f64_noalloc": (_c_int, [_c_int, _vp, _vp, _vp]),
    "pt_verlet_positions": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _c_dbl, _vp, _vp]),
    "pt_verlet_velocities": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _c_dbl, _vp]),
    "pt_kinetic_energy": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _c_size, _vp]),
    "pt_berendsen_rescale": (_c_int, [_c_int, _vp, _vp, _c_dbl, _c_dbl, _c_dbl, _c_dbl, _vp]),
    "pt_gather_rows_f64": (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _vp]),
    "pt_scatter_rows_f64": (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _vp]),
    "pt_gather_rows_i32": (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _vp]),
    "pt_wrap_positions": (_c_int, [_c_int, _vp, _vp, _vp]),
    "pt_max_displacement": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp]),
    "pt_mlp_param_count": (_c_size, [_vp]),
    "pt_mlp_train_grad": (_c_int, [_c_int, _c_int, _vp, _vp, _c_int, _c_int, _vp, _vp, _vp, _vp,
                                   _c_dbl, _vp, _vp]),
    "pt_train_seeds": (_c_int, [_c_int, _vp, _vp, _vp, _c_dbl, _c_dbl, _c_dbl, _vp, _vp, _vp, _vp]),
    "pt_adam_step": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _c_dbl, _c_dbl, _c_dbl, _c_dbl, _c_dbl,
                              _c_int, _vp]),
    "pt_mlp_jacobian": (_c_int, [_c_int, _c_int, _vp, _vp, _c_ll, _c_int, _vp, _vp, _vp, _c_dbl, _vp, _c_ll, _vp]),
    "pt_kalman_residual": (_c_int, [_c_int, _vp, _vp, _c_dbl, _c_dbl, _vp, _vp, _vp]),
    "pt_kalman_workspace_bytes": (_c_size, [_c_int, _c_int]),
    "pt_kalman_update": (_c_int, [_c_int, _c_int, _vp, _vp, _c_ll, _vp, _vp, _c_dbl, _c_dbl, _c_dbl, _vp, _c_size,
                                  _vp, _vp]),
    "pt_permute_state_tmp_bytes": (_c_size, [_c_int, _c_int, _vp]),
    "pt_permute_state_gated": (_c_int, [_c_int, _vp, _c_int, _vp, _vp, _c_size, _vp, _vp]),
}

#: Every symbol the header declares; the CPU test suite checks each is exported.
EXPORTED_SYMBOLS = tuple(_SIGS)

_lib: Optional[ctypes.CDLL] = None


def load() -> ctypes.CDLL:
    """Load the library (raises HipLibraryError when it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise HipLibraryError(
            f"{LIB_PATH} is missing: build it with `python -m pantea_amd.build` "
            "(the GPU path has no CPU fallback)"
        )
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    msg = load().pt_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str) -> None:
    if rc != PT_OK:
        raise HipLibraryError(f"{what} failed ({rc}): {last_error()}")


def require_device() -> "object":
    """Return the torch CUDA/HIP device; raise when none is present."""
    import torch

    load()
    if not torch.cuda.is_available():
        raise HipLibraryError("no HIP device visible: pantea_amd runs on the GPU only")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr() -> int:
    import torch

    return torch.cuda.current_stream().cuda_stream


def ptr(t) -> Optional[int]:
    """Device pointer of a tensor (None passes NULL)."""
    if t is None:
        return None
    return t.data_ptr()
