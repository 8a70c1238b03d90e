"""ctypes binding of ``libr2splat_hip.so`` (the C ABI in include/r2splat_hip.h).

The library is loaded lazily on first use. When it is missing, or when no GPU
is visible to PyTorch, every operator raises RuntimeError: there is no CPU
fallback. Mirrors the lazy loader of r2splat/cuda/_backend.py:16-45.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import torch

LIB_PATH = Path(os.environ.get("R2S_LIBRARY", Path(__file__).resolve().parent.parent / "libr2splat_hip.so"))

F32, F64 = 0, 1

P = ctypes.c_void_p
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64
I32 = ctypes.c_int
F = ctypes.c_float
D = ctypes.c_double
SZ = ctypes.c_size_t


class ValueCheck(ctypes.Structure):
    """One entry of an r2s_value_checks batch (r2s_value_check)."""

    _fields_ = [("data", P), ("n", U64), ("dtype", ctypes.c_int32), ("kind", ctypes.c_int32), ("row", U32),
                ("reserved", U32)]


CHECK_NONFINITE, CHECK_NONPOSITIVE, CHECK_NEGATIVE, CHECK_ZERO_ROW = 0, 1, 2, 3
CHECK_K_FOCAL, CHECK_K_PINHOLE, CHECK_VIEW_ROW = 4, 5, 6
MAX_VALUE_CHECKS = 16


class RasterJob(ctypes.Structure):
    """One entry of a packed render / raster-VJP table (r2s_raster_job)."""

    _fields_ = [("packed", P), ("tile_offsets", P), ("tile_order", P), ("counters", P), ("image", P),
                ("image_var", P), ("v_image", P), ("v_image_var", P), ("isect_grads", P), ("capacity", U32),
                ("C", U32), ("width", U32), ("height", U32), ("detach", ctypes.c_int32), ("first_tile", U32),
                ("reserved", U32 * 2)]


# name -> argument types (return type is int status unless listed in _RESTYPES)
SIGNATURES = {
    "r2s_abi_version": [],
    "r2s_last_error": [],
    "r2s_value_checks": [P, U32, P, P],
    "r2s_parameter_transform_fwd": [I32, P, P, P, P, U32, P],
    "r2s_parameter_transform_bwd": [I32, P, P, P, P, P, U32, P],
    "r2s_world_to_image_projection_fwd": [I32, P, P, P, P, U32, U32, F, F, P, P, P, U32, U32, P],
    "r2s_world_to_image_projection_bwd": [I32, P, P, P, P, U32, U32, P, P, P, P, U32, U32, P],
    "r2s_conics_mu_fwd": [I32, P, D, P, P, U32, P],
    "r2s_conics_mu_bwd": [I32, P, D, P, P, P, P, P, U32, P],
    "r2s_world_to_volume_mapping_fwd": [I32, P, P, P, P, P, U32, U32, U32, D, D, P, P, P, U32, P],
    "r2s_world_to_volume_mapping_bwd": [I32, P, P, P, P, U32, U32, U32, D, P, P, P, P, U32, P],
    "r2s_scan_workspace_bytes": [U64],
    "r2s_sort_workspace_bytes": [U64, U32],
    "r2s_isect_tiles_count": [I32, P, P, I32, U32, U32, U32, U32, U32, P, P, P, P, P],
    "r2s_isect_tiles_emit": [I32, P, P, I32, U32, U32, U32, U32, U32, P, U32, P, P, P, SZ, P],
    "r2s_isect_tiles_3d_count": [I32, P, P, U32, U32, U32, U32, U32, P, P, P, P, P],
    "r2s_isect_tiles_3d_emit": [I32, P, P, U32, U32, U32, U32, U32, P, U32, P, P, P, SZ, P],
    "r2s_isect_offset_encode": [P, U32, U32, P, P],
    "r2s_rasterize_to_pixels_fwd": [I32, P, P, P, P, P, P, U32, U32, U32, U32, U32, U32, U32, U32, P, P],
    "r2s_rasterize_to_pixels_mean_var_fwd": [I32, P, P, P, P, P, P, P, U32, U32, U32, U32, U32, U32, U32, U32, P, P, P],
    "r2s_beta_rasterize_to_pixels_fwd": [I32, P, P, P, P, P, P, P, U32, U32, U32, U32, U32, U32, U32, U32, P, P],
    "r2s_beta_rasterize_to_pixels_mean_var_fwd": [I32, P, P, P, P, P, P, P, P, U32, U32, U32, U32, U32, U32, U32, U32, P, P, P],
    "r2s_rasterize_to_pixels_uncertainty_bwd": [I32, P, P, P, P, P, P, P, U32, U32, U32, U32, U32, U32, U32, U32, P, P, P],
    "r2s_rasterize_to_pixels_bwd": [I32, P, P, P, P, P, P, P, U32, U32, U32, U32, U32, U32, U32, U32, P, P, P, P, P, P, P, P],
    "r2s_rasterize_to_pixels_bwd_per_gaussian": [I32, P, P, P, P, P, P, P, U32, U32, U32, U32, U32, U32, U32, U32, P, P, P, P, P, P, P, P, P, P, P],
    "r2s_voxelize_to_voxels_fwd": [I32, P, P, P, P, P, P, P, U32, U32, U32, U32, U32, U32, U32, U32, U32, P, P, P],
    "r2s_voxelize_to_voxels_bwd": [I32, P, P, P, P, P, P, U32, U32, U32, U32, U32, U32, U32, U32, U32, P, P, P, P, P, P],
    "r2s_voxelize_to_voxels_bwd_per_gaussian": [I32, P, P, P, P, P, P, U32, U32, U32, U32, U32, U32, U32, U32, U32, P, P, P, P, P, P, P, P, P],
    "r2s_fused_scratch_bytes": [U32, U32, U32, U32, U32, U32, I32],
    "r2s_fused_record_bytes": [I32],
    "r2s_fused_packed_bytes": [U32, U32, U32, I32],
    "r2s_fused_forward": [P, P, P, P, P, P, P, P, P, U32, U32, U32, U32, U32, F, F, U32, P, SZ, P, P, P, P, P, P, P,
                          P],
    "r2s_fused_backward": [P, P, P, P, P, P, P, P, P, P, U32, U32, U32, U32, U32, F, U32, P, P, P, P, P, P, P, SZ, P,
                           P, I32, P, P, P, P, P, P, P],
    "r2s_fused_prepare": [P, P, P, P, P, P, P, P, P, U32, U32, U32, U32, U32, F, F, U32, P, SZ, P, P, P, P, P, P],
    "r2s_fused_render": [P, P, P, P, U32, U32, U32, U32, U32, U32, I32, P, P, P],
    "r2s_fused_raster_vjp": [P, P, P, P, U32, U32, U32, U32, U32, U32, I32, P, P, I32, P, P],
    "r2s_fused_render_packed": [P, P, U32, U32, U32, I32, P],
    "r2s_fused_raster_vjp_packed": [P, P, U32, U32, U32, I32, P],
    "r2s_fused_keep_masks": [P, P, P, U32, U32, U32, U32, U32, I32, P, P, P],
    "r2s_fused_param_workspace_bytes": [U32, U32],
    "r2s_fused_param_vjp": [P, P, P, P, P, P, P, P, P, U32, U32, U32, U32, F, P, P, U32, P, P, SZ, P, P, P, P, P, P,
                            P],
    "r2s_voxel_frame": [P, P, P, P, P, U32, U32, U32, D, U32, P, P],
    "r2s_voxelize_fused_fwd": [P, P, P, P, P, P, P, P, U32, U32, U32, U32, U32, U32, U32, U32, U32, P, P, P],
    "r2s_voxel_partials_bytes": [U32],
    "r2s_voxelize_fused_vjp": [P, P, P, P, P, P, P, P, P, P, U32, U32, U32, U32, U32, U32, U32, U32, U32, P,
                               P, I32, P, SZ, P, P, P, P, P, P],
    "r2s_mse_partials": [U64],
    "r2s_mse_adjoint": [P, P, U64, P, P, P, P],
    "r2s_mse_adjoint_noalloc": [P, P, U64, P, P, P],
    "r2s_adam_step": [P, P, P, P, U64, F, F, F, F, P, P, P],
    "r2s_clamp_min": [P, U64, F, P],
    "r2s_volume_tv_fwd": [I32, P, U32, U32, U32, U32, P, P, P],
    "r2s_volume_tv_bwd": [I32, P, P, U32, U32, U32, U32, F, P, P],
    "r2s_volume_tv_iso_fwd": [I32, P, U32, U32, U32, U32, D, P, P, P],
    "r2s_volume_tv_iso_bwd": [I32, P, P, U32, U32, U32, U32, D, D, P, P],
    "r2s_image_l1_fwd": [I32, P, P, U64, P, P, P],
    "r2s_image_l1_bwd": [I32, P, P, P, U64, F, P, P],
}

_RESTYPES = {
    "r2s_last_error": ctypes.c_char_p,
    "r2s_scan_workspace_bytes": SZ,
    "r2s_sort_workspace_bytes": SZ,
    "r2s_fused_scratch_bytes": SZ,
    "r2s_fused_record_bytes": SZ,
    "r2s_fused_packed_bytes": SZ,
    "r2s_fused_param_workspace_bytes": SZ,
    "r2s_voxel_partials_bytes": SZ,
    "r2s_mse_partials": SZ,
}


class _Library:
    def __init__(self):
        self._lib = None
        self._error = None
        self._gpu = False

    def load(self, require_gpu: bool = True):
        if self._gpu:
            return self._lib
        if self._lib is None and self._error is None:
            try:
                lib = ctypes.CDLL(str(LIB_PATH))
                for name, argtypes in SIGNATURES.items():
                    fn = getattr(lib, name)
                    fn.argtypes = argtypes
                    fn.restype = _RESTYPES.get(name, ctypes.c_int)
                self._lib = lib
            except OSError as error:
                self._error = f"cannot load {LIB_PATH}: {error}"
        if self._lib is None:
            raise RuntimeError(
                f"r2splat_amd HIP library is not available ({self._error}). "
                "Build it with `python -m r2splat_amd._build`."
            )
        if require_gpu:
            if not torch.cuda.is_available():
                raise RuntimeError("r2splat_amd operators require a ROCm GPU; none is visible to PyTorch")
            self._gpu = True
        return self._lib


_LIBRARY = _Library()


def lib(require_gpu: bool = True):
    """The loaded library (raises RuntimeError when unavailable)."""
    return _LIBRARY.load(require_gpu)


def library_path() -> Path:
    return LIB_PATH


def is_available() -> bool:
    try:
        lib(require_gpu=True)
        return True
    except RuntimeError:
        return False


def call(name: str, *args) -> None:
    """Calls a status-returning entry point, raising RuntimeError on failure."""
    handle = lib()
    status = getattr(handle, name)(*args)
    if status != 0:
        message = handle.r2s_last_error().decode(errors="replace")
        raise RuntimeError(f"{name}: {message}")


def dtype_code(dtype: torch.dtype) -> int:
    if dtype == torch.float32:
        return F32
    if dtype == torch.float64:
        return F64
    raise RuntimeError(f"unsupported dtype {dtype}; expected float32 or float64")


def stream_of(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t) -> int | None:
    return None if t is None else t.data_ptr()
