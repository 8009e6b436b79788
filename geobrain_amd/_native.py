"""ctypes binding of ``libgbw_acoustic2d.so`` (declared in ``include/gbw_acoustic2d.h``
and ``include/gbw_acoustic3d.h``).

The library is the only compute path: if it is missing, or no ROCm GPU is
visible, calls raise ``RuntimeError``.  There is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgbw_acoustic2d.so")

c_i32 = ctypes.c_int32
c_f32 = ctypes.c_float
c_vp = ctypes.c_void_p

ABI_VERSION = 6
# size of the device array behind gbw_acoustic2d_forward_noalloc (GBW_NOALLOC_SCRATCH_BYTES)
NOALLOC_SCRATCH_BYTES = 64 << 20

EXPORTS = (
    "gbw_abi_version",
    "gbw_last_error",
    "gbw_acoustic2d_workspace_bytes",
    "gbw_acoustic2d_checkpoint_bytes",
    "gbw_acoustic2d_forward",
    "gbw_acoustic2d_forward_noalloc",
    "gbw_acoustic2d_geometry_bytes",
    "gbw_acoustic2d_geometry_setup",
    "gbw_acoustic2d_backward",
    "gbw_set_kernel_timing",
    "gbw_kernel_times",
    "gbw_acoustic2d_batch_workspace_bytes",
    "gbw_acoustic2d_batch_setup",
    "gbw_acoustic2d_batch_forward",
    "gbw_device_wall_clock_khz",
    "gbw_acoustic2d_batch_gradient_workspace_bytes",
    "gbw_acoustic2d_batch_gradient_setup",
    "gbw_acoustic2d_batch_gradient_forward",
    "gbw_acoustic2d_batch_gradient_backward",
    "gbw_acoustic3d_workspace_bytes",
    "gbw_acoustic3d_checkpoint_bytes",
    "gbw_acoustic3d_forward",
    "gbw_acoustic3d_backward",
)


class Problem(ctypes.Structure):
    """``gbw_acoustic2d_problem``."""
    _fields_ = [
        ("nx", c_i32), ("nz", c_i32), ("dx", c_f32), ("dz", c_f32), ("nt", c_i32),
        ("dt", c_f32), ("nabc", c_i32), ("free_surface", c_i32), ("src_n", c_i32),
        ("rcv_n", c_i32), ("wavelet_shared", c_i32), ("energy_chunk", c_i32),
        ("snapshot_every", c_i32), ("checkpoint_every", c_i32), ("checkpoint_bf16", c_i32),
        ("tile_rows", c_i32), ("shots_per_block", c_i32), ("march", c_i32),
        ("steps_per_launch", c_i32), ("fwd_march", c_i32), ("src_points", c_i32),
    ]


class Inputs(ctypes.Structure):
    """``gbw_acoustic2d_inputs``."""
    _fields_ = [
        ("src_x", c_vp), ("src_z", c_vp), ("rcv_x", c_vp), ("rcv_z", c_vp),
        ("src_v", c_vp), ("damp", c_vp), ("vp", c_vp), ("rho", c_vp), ("geometry", c_vp),
    ]


class Outputs(ctypes.Structure):
    """``gbw_acoustic2d_outputs``."""
    _fields_ = [
        ("rec_p", c_vp), ("rec_vx", c_vp), ("rec_vz", c_vp),
        ("energy_p", c_vp), ("energy_vx", c_vp), ("energy_vz", c_vp),
        ("snapshots", c_vp), ("checkpoints", c_vp),
    ]


class Adjoint(ctypes.Structure):
    """``gbw_acoustic2d_adjoint``."""
    _fields_ = [
        ("g_p", c_vp), ("g_vx", c_vp), ("g_vz", c_vp), ("checkpoints", c_vp),
        ("grad_vp", c_vp), ("grad_rho", c_vp), ("grad_src_v", c_vp),
    ]


class Problem3(ctypes.Structure):
    """``gbw_acoustic3d_problem``."""
    _fields_ = [
        ("nx", c_i32), ("ny", c_i32), ("nz", c_i32), ("dx", c_f32), ("dy", c_f32),
        ("dz", c_f32), ("nt", c_i32), ("dt", c_f32), ("nabc", c_i32), ("free_surface", c_i32),
        ("src_n", c_i32), ("rcv_n", c_i32), ("wavelet_shared", c_i32), ("energy_chunk", c_i32),
        ("shots_per_block", c_i32), ("checkpoint_every", c_i32), ("step_kernel", c_i32),
        ("z_chunk", c_i32), ("src_points", c_i32),
    ]


class Inputs3(ctypes.Structure):
    """``gbw_acoustic3d_inputs``."""
    _fields_ = [
        ("src_x", c_vp), ("src_y", c_vp), ("src_z", c_vp), ("rcv_x", c_vp), ("rcv_y", c_vp),
        ("rcv_z", c_vp), ("src_v", c_vp), ("damp", c_vp), ("vp", c_vp), ("rho", c_vp),
    ]


class Outputs3(ctypes.Structure):
    """``gbw_acoustic3d_outputs``."""
    _fields_ = [
        ("rec_p", c_vp), ("rec_vx", c_vp), ("rec_vy", c_vp), ("rec_vz", c_vp),
        ("energy_p", c_vp), ("energy_vx", c_vp), ("energy_vy", c_vp), ("energy_vz", c_vp),
        ("checkpoints", c_vp),
    ]


class Adjoint3(ctypes.Structure):
    """``gbw_acoustic3d_adjoint``."""
    _fields_ = [
        ("g_p", c_vp), ("g_vx", c_vp), ("g_vy", c_vp), ("g_vz", c_vp), ("checkpoints", c_vp),
        ("grad_vp", c_vp), ("grad_rho", c_vp), ("grad_src_v", c_vp),
    ]


_lock = threading.Lock()
_lib = None

STATUS = {
    -1: ValueError,
    -2: IndexError,
    -3: RuntimeError,
    -4: RuntimeError,
    -5: NotImplementedError,
}


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the library once; raises ``RuntimeError`` if it was not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise RuntimeError(
                f"geobrain_amd: HIP library not found at {path}; build it with "
                "`python -m geobrain_amd.build` (there is no CPU fallback)")
        # torch bundles its own libamdhip64.so.7; loading torch first makes the
        # library bind to that same HIP runtime (and so share torch's device
        # context and streams) instead of pulling in a second copy from /opt/rocm
        import torch  # noqa: F401
        lib = ctypes.CDLL(path)
        lib.gbw_abi_version.restype = ctypes.c_int
        lib.gbw_abi_version.argtypes = []
        lib.gbw_last_error.restype = ctypes.c_char_p
        lib.gbw_last_error.argtypes = []
        lib.gbw_acoustic2d_workspace_bytes.restype = ctypes.c_size_t
        lib.gbw_acoustic2d_workspace_bytes.argtypes = [ctypes.POINTER(Problem), ctypes.c_int]
        lib.gbw_acoustic2d_checkpoint_bytes.restype = ctypes.c_size_t
        lib.gbw_acoustic2d_checkpoint_bytes.argtypes = [ctypes.POINTER(Problem)]
        lib.gbw_acoustic2d_forward.restype = ctypes.c_int
        lib.gbw_acoustic2d_forward.argtypes = [
            ctypes.POINTER(Problem), ctypes.POINTER(Inputs), ctypes.POINTER(Outputs),
            c_vp, ctypes.c_size_t, c_vp]
        lib.gbw_acoustic2d_geometry_bytes.restype = ctypes.c_size_t
        lib.gbw_acoustic2d_geometry_bytes.argtypes = [ctypes.POINTER(Problem)]
        lib.gbw_acoustic2d_geometry_setup.restype = ctypes.c_int
        lib.gbw_acoustic2d_geometry_setup.argtypes = [
            ctypes.POINTER(Problem), ctypes.POINTER(Inputs), c_vp, ctypes.c_size_t, c_vp]
        lib.gbw_acoustic2d_forward_noalloc.restype = ctypes.c_int
        lib.gbw_acoustic2d_forward_noalloc.argtypes = [
            ctypes.POINTER(Problem), ctypes.POINTER(Inputs), ctypes.POINTER(Outputs), c_vp]
        lib.gbw_acoustic2d_backward.restype = ctypes.c_int
        lib.gbw_acoustic2d_backward.argtypes = [
            ctypes.POINTER(Problem), ctypes.POINTER(Inputs), ctypes.POINTER(Adjoint),
            c_vp, ctypes.c_size_t, c_vp]
        lib.gbw_set_kernel_timing.restype = None
        lib.gbw_set_kernel_timing.argtypes = [ctypes.c_int]
        lib.gbw_kernel_times.restype = ctypes.c_size_t
        lib.gbw_kernel_times.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
        lib.gbw_acoustic2d_batch_workspace_bytes.restype = ctypes.c_size_t
        lib.gbw_acoustic2d_batch_workspace_bytes.argtypes = [ctypes.POINTER(Problem), ctypes.c_int]
        lib.gbw_acoustic2d_batch_setup.restype = ctypes.c_int
        lib.gbw_acoustic2d_batch_setup.argtypes = [
            ctypes.POINTER(Problem), ctypes.POINTER(Inputs), ctypes.POINTER(Outputs), ctypes.c_int,
            c_vp, ctypes.c_size_t, c_vp]
        lib.gbw_acoustic2d_batch_forward.restype = ctypes.c_int
        lib.gbw_acoustic2d_batch_forward.argtypes = [
            ctypes.POINTER(Problem), ctypes.c_int, c_vp, ctypes.c_size_t, c_vp, c_vp]
        lib.gbw_acoustic2d_batch_gradient_workspace_bytes.restype = ctypes.c_size_t
        lib.gbw_acoustic2d_batch_gradient_workspace_bytes.argtypes = [
            ctypes.POINTER(Problem), ctypes.c_int]
        lib.gbw_acoustic2d_batch_gradient_setup.restype = ctypes.c_int
        lib.gbw_acoustic2d_batch_gradient_setup.argtypes = [
            ctypes.POINTER(Problem), ctypes.POINTER(Inputs), ctypes.POINTER(Outputs),
            ctypes.POINTER(Adjoint), ctypes.c_int, c_vp, ctypes.c_size_t, c_vp]
        for name in ("gbw_acoustic2d_batch_gradient_forward", "gbw_acoustic2d_batch_gradient_backward"):
            fn = getattr(lib, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.POINTER(Problem), ctypes.c_int, c_vp, ctypes.c_size_t, c_vp, c_vp]
        lib.gbw_device_wall_clock_khz.restype = ctypes.c_int
        lib.gbw_device_wall_clock_khz.argtypes = []
        lib.gbw_acoustic3d_workspace_bytes.restype = ctypes.c_size_t
        lib.gbw_acoustic3d_workspace_bytes.argtypes = [ctypes.POINTER(Problem3), ctypes.c_int]
        lib.gbw_acoustic3d_checkpoint_bytes.restype = ctypes.c_size_t
        lib.gbw_acoustic3d_checkpoint_bytes.argtypes = [ctypes.POINTER(Problem3)]
        lib.gbw_acoustic3d_forward.restype = ctypes.c_int
        lib.gbw_acoustic3d_forward.argtypes = [
            ctypes.POINTER(Problem3), ctypes.POINTER(Inputs3), ctypes.POINTER(Outputs3),
            c_vp, ctypes.c_size_t, c_vp]
        lib.gbw_acoustic3d_backward.restype = ctypes.c_int
        lib.gbw_acoustic3d_backward.argtypes = [
            ctypes.POINTER(Problem3), ctypes.POINTER(Inputs3), ctypes.POINTER(Adjoint3),
            c_vp, ctypes.c_size_t, c_vp]
        ver = lib.gbw_abi_version()
        if ver != ABI_VERSION:
            raise RuntimeError(f"geobrain_amd: ABI version {ver} != expected {ABI_VERSION}")
        _lib = lib
        return lib


def check(rc: int) -> None:
    """Raise the Python exception matching a non-zero gbw_status."""
    if rc == 0:
        return
    msg = load().gbw_last_error().decode(errors="replace")
    raise STATUS.get(rc, RuntimeError)(msg)


def last_error() -> str:
    return load().gbw_last_error().decode(errors="replace")


def set_kernel_timing(on: bool) -> None:
    """Enable (and reset) or disable the library's per-kernel event timing."""
    load().gbw_set_kernel_timing(int(bool(on)))


def kernel_times() -> dict:
    """Per-kernel ``{name: (launches, total_ms)}`` since timing was enabled."""
    lib = load()
    n = lib.gbw_kernel_times(None, 0)
    buf = ctypes.create_string_buffer(n)
    lib.gbw_kernel_times(buf, n)
    out = {}
    for line in buf.value.decode().splitlines():
        name, launches, ms = line.rsplit(" ", 2)
        out[name] = (int(launches), float(ms))
    return out
