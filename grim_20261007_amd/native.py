"""ctypes binding of the GRIM-2D HIP library (include/grim2d.h).

The library is loaded from this package directory; a missing or stale build
raises at import of the product path (there is no CPU fallback).  Device
memory and streams come from PyTorch-ROCm: tensors are passed as raw device
pointers, streams as ``torch.cuda.current_stream().cuda_stream``.
"""
from __future__ import annotations

import ctypes
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libgrim2d.so"

_i64 = ctypes.c_int64
_i32 = ctypes.c_int
_dbl = ctypes.c_double
_vp = ctypes.c_void_p

SIGNATURES = {
    "grim_version": (_i32, []),
    "grim_init": (_i32, [_i32]),
    "grim_assemble_workspace_bytes": (_i64, [_i64, _i64]),
    "grim_assemble": (_i32, [_i64, _vp, _vp, _vp, _vp, _vp, _i64, _dbl, _i32, _vp, _i64, _vp, _i64, _vp, _i64,
                             _i64, _vp]),
    "grim_assemble_far": (_i32, [_i64, _vp, _vp, _vp, _vp, _vp, _i64, _dbl, _i32, _vp, _i64, _vp, _i64, _vp,
                                 _i64, _i64, _vp]),
    "grim_assemble_near": (_i32, [_i64, _vp, _vp, _vp, _vp, _vp, _i64, _dbl, _i32, _vp, _i64, _vp, _i64, _vp,
                                  _i64, _i64, _vp]),
    "grim_assemble_noalloc": (_i32, [_i64, _vp, _vp, _vp, _vp, _vp, _i64, _dbl, _i32, _vp, _i64, _vp, _i64, _i64,
                                     _vp, _vp]),
    "grim_workspace_status": (_i32, [_vp, _vp]),
    "grim_loads": (_i32, [_i64, _i64, _vp, _vp, _vp, _dbl, _vp, _vp, _vp, _i64, _vp]),
    "grim_farfield": (_i32, [_i64, _i64, _vp, _i64, _vp, _vp, _i64, _dbl, _vp, _vp, _vp]),
    "grim_farfield_potential": (_i32, [_i64, _i64, _vp, _i64, _vp, _i64, _dbl, _vp, _vp, _vp]),
    "grim_assemble_operators_lossy_workspace_bytes": (_i64, [_i64, _i64]),
    "grim_assemble_operators_lossy": (_i32, [_i64, _vp, _vp, _vp, _vp, _dbl, _dbl, _i32, _vp, _vp, _vp, _vp, _i64,
                                             _dbl, _dbl, _vp, _vp, _vp, _i64, _i64, _vp]),
    "grim_hankel2_probe": (_i32, [_i64, _vp, _vp, _vp]),
    "grim_sheet_workspace_bytes": (_i64, [_i64, _i64]),
    "grim_assemble_sheet": (_i32, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _dbl, _i32, _i32, _vp, _vp, _vp,
                                   _vp, _i64, _vp, _i64, _vp, _i64, _i64, _vp]),
    "grim_loads_sheet": (_i32, [_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _dbl, _vp, _vp, _vp, _i64, _vp]),
    "grim_lu_aux_bytes": (_i64, [_i64, _i32]),
    "grim_lu_factor": (_i32, [_i64, _i64, _vp, _i64, _i64, _i32, _vp, _i64, _vp, _i64, _vp, _vp]),
    "grim_lu_solve_workspace_bytes": (_i64, [_i64, _i64, _i64, _i32]),
    "grim_lu_solve": (_i32, [_i64, _i64, _i64, _vp, _i64, _i64, _i32, _vp, _i64, _vp, _i64, _i64, _vp, _i64, _vp]),
    "grim_gemm_minus": (_i32, [_i64, _i64, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _i32,
                               _vp]),
    "grim_refine_workspace_bytes": (_i64, [_i64, _i64, _i64]),
    "grim_refined_solve": (_i32, [_i64, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _vp, _i64, _i64,
                                  _vp, _i64, _i64, _i32, _dbl, _vp, _i64, ctypes.POINTER(_i32),
                                  ctypes.POINTER(_dbl), _vp]),
    "grim_lu_solve_ex_workspace_bytes": (_i64, [_i64, _i64, _i64, _i32, _i32]),
    "grim_lu_solve_ex": (_i32, [_i64, _i64, _i64, _vp, _i64, _i64, _i32, _vp, _i64, _i32, _vp, _i64, _i64, _vp,
                                _i64, _vp]),
    "grim_gemm_minus_ex": (_i32, [_i64, _i64, _i64, _i64, _i32, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64,
                                  _i32, _vp]),
    "grim_equilibration": (_i32, [_i64, _i64, _vp, _i64, _i64, _vp, _vp, _vp, _vp]),
    "grim_norm_inf_workspace_bytes": (_i64, [_i64, _i64]),
    "grim_norm_inf": (_i32, [_i64, _i64, _vp, _i64, _i64, _vp, _i64, _vp, _vp]),
    "grim_refine_ex_workspace_bytes": (_i64, [_i64, _i64, _i64, _i32]),
    "grim_refined_solve_ex": (_i32, [_i64, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i32, _vp, _i64,
                                     _i64, _vp, _i64, _i64, _i32, _dbl, _vp, _i64, ctypes.POINTER(_i32),
                                     ctypes.POINTER(_dbl), _vp]),
    "grim_to_c64": (_i32, [_i64, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _vp]),
    "grim_copy_c128": (_i32, [_i64, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _vp]),
    "grim_column_norms": (_i32, [_i64, _i64, _vp, _i64, _vp, _vp]),
    "grim_robin_rows": (_i32, [_i64, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _vp, _vp]),
    "grim_assemble_operators": (_i32, [_i64, _vp, _vp, _vp, _vp, _vp, _i64, _dbl, _i32, _i32, _vp, _i64, _vp, _i64,
                                        _dbl, _vp, _vp, _i64, _i64, _vp]),
    "grim_route": (_i32, [_i64, _i64, _vp, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp]),
    "grim_scale_blocks": (_i32, [_i64, _vp, _i64, _vp, _vp]),
    "grim_robin_split": (_i32, [_i64, _vp, _vp, _vp, _vp, _dbl, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _i64, _i64,
                                 _vp, _i64, _vp, _i64, _vp]),
    "grim_robin_rhs": (_i32, [_i64, _i64, _vp, _vp, _vp, _dbl, _vp, _vp, _vp, _vp, _i64, _vp]),
    "grim_assemble_poly_workspace_bytes": (_i64, [_i64, _i32, _i64]),
    "grim_assemble_poly": (_i32, [_i64, _i32, _vp, _vp, _vp, _vp, _dbl, _i32, _vp, _i64, _vp, _i64, _vp, _i64, _i64,
                                  _vp]),
    "grim_loads_poly": (_i32, [_i64, _i32, _i64, _vp, _vp, _vp, _dbl, _vp, _i64, _vp, _vp, _vp, _vp, _i64, _vp]),
    "grim_set_launch_shape": (_i32, [_i32, _i64, _i32]),
    "grim_clear_launch_shapes": (_i32, []),
    "grim_launch_shape": (_i32, [_i32, _i64]),
}

STATUS = {0: "ok", 1: "size or index out of range", 3: "invalid arguments", 5: "HIP runtime failure",
          6: "not converged", 7: "near-pair capacity exceeded", 8: "zero pivot",
          10: "coincident integration points"}


class GrimError(RuntimeError):
    def __init__(self, func: str, code: int):
        super().__init__(f"{func} failed with status {code} ({STATUS.get(code, 'unknown')})")
        self.code = code


_lib = None
_lock = threading.Lock()
_initialised = set()


def load(path: Path | str = LIB_PATH) -> ctypes.CDLL:
    """Load the shared library (raises OSError if it was not built)."""
    global _lib
    # PyTorch's HIP runtime is loaded first, so that the library's libamdhip64
    # dependency binds to that same runtime (one runtime, one set of streams).
    import torch  # noqa: F401
    with _lock:
        if _lib is None:
            p = Path(path)
            if not p.exists():
                raise OSError(f"GRIM-2D HIP library not built: {p} (run python -m grim_20261007_amd.build)")
            lib = ctypes.CDLL(str(p))
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def lib() -> ctypes.CDLL:
    return _lib if _lib is not None else load()


def check(name: str, code: int) -> None:
    if code != 0:
        raise GrimError(name, code)


def ensure_device(device: int) -> None:
    """grim_init once per device (uploads quadrature rules to constant memory).

    PyTorch's HIP runtime is initialised first: it must own the device context
    that the library's launches then share."""
    with _lock:
        if device in _initialised:
            return
    import torch
    torch.cuda.init()
    check("grim_init", lib().grim_init(int(device)))
    with _lock:
        _initialised.add(device)


def ptr(t) -> int:
    return t.data_ptr() if t is not None else None
