"""ctypes binding of libkarman_hip.so (the C-ABI declared in include/karman.h).

The shared library is the only compute path: if it is missing or no HIP
device is available, ``lib()`` / ``Context`` raise instead of falling back to
anything else.
"""
from __future__ import annotations

import ctypes
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libkarman_hip.so"

KM_OK, KM_BREAKDOWN, KM_MAXITER, KM_NONFINITE, KM_STALL = 0, 1, 2, 3, 4
KM_EINVAL, KM_EHIP, KM_ENOMEM = -1, -2, -3
KM_F64, KM_F32 = 0, 1
NOALLOC_MAX_CELLS = 521144    # KM_NOALLOC_MAX_CELLS: largest system km_bicgstab_noalloc takes
KM_FLAG_X0, KM_FLAG_MIXED, KM_FLAG_NO_GRAPH, KM_FLAG_COARSE, KM_FLAG_SMOOTH, KM_FLAG_TRUE = 1, 2, 4, 8, 16, 32
KM_FLAG_GCR, KM_FLAG_JACOBI, KM_FLAG_MG, KM_FLAG_SYMMETRIC = 64, 128, 256, 512
KM_EAST_WALL, KM_EAST_OUTFLOW = 0, 1
KM_KERNEL_BICG_A, KM_KERNEL_BICG_B, KM_KERNEL_BICG_R, KM_KERNEL_STENCIL = 0, 1, 2, 3
KM_KERNEL_TL_PV, KM_KERNEL_TL_CV, KM_KERNEL_TL_ST, KM_KERNEL_TL_CT = 4, 5, 6, 7
KM_KERNEL_BACK_TO_BACK, KM_KERNEL_SYMMETRIC = 0x100, 0x200

STATUS_NAMES = {KM_OK: "ok", KM_BREAKDOWN: "breakdown", KM_MAXITER: "maxiter",
                KM_NONFINITE: "nonfinite", KM_STALL: "stall"}


class KmGrid(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("east", ctypes.c_int32), ("pin_pcor", ctypes.c_int32),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("rho", ctypes.c_double),
        ("mu", ctypes.c_double), ("dt", ctypes.c_double),
        ("alpha_u", ctypes.c_double), ("alpha_p", ctypes.c_double), ("alpha_m", ctypes.c_double),
        ("bc", (ctypes.c_double * 2) * 4),
    ]


class KmSolveInfo(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32), ("status", ctypes.c_int32),
        ("refinements", ctypes.c_int32), ("fallback", ctypes.c_int32),
        ("rel_resid", ctypes.c_double), ("rel_resid_true", ctypes.c_double),
        ("bnorm", ctypes.c_double),
    ]


class KmCase(ctypes.Structure):
    """One case of a batched step (include/karman.h km_case)."""
    _fields_ = [("g", KmGrid)] + [(f, ctypes.c_void_p) for f in (
        "u", "v", "p", "u0", "v0", "u_mid", "v_mid", "pcor", "mdiv",
        "coef_u", "coef_v", "coef_p", "b_u", "b_v", "b_p", "mask")]


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_CTX = ctypes.c_void_p
_GRID = ctypes.POINTER(KmGrid)
_INFO = ctypes.POINTER(KmSolveInfo)

# name -> argument types (every function returns int unless listed in _RESTYPE)
SIGNATURES = {
    "km_create": [ctypes.POINTER(ctypes.c_void_p), _I, _P],
    "km_destroy": [_CTX],
    "km_set_stream": [_CTX, _P],
    "km_last_error": [_CTX],
    "km_synchronize": [_CTX],
    "km_set_tile": [_CTX, _I, _I, _I, _I, _I],
    "km_get_tile": [_CTX, _I, _I, _I, ctypes.POINTER(_I), ctypes.POINTER(_I)],
    "km_set_autotune": [_CTX, _I],
    "km_assemble_u": [_CTX, _GRID, _P, _P, _P, _P, _P, _P, _P],
    "km_assemble_v": [_CTX, _GRID, _P, _P, _P, _P, _P, _P, _P],
    "km_apply_ghosts": [_CTX, _GRID, _P, _P],
    "km_blend": [_CTX, _GRID, _P, _P, _P, _P],
    "km_outflow_fix": [_CTX, _GRID, _P],
    "km_assemble_p": [_CTX, _GRID, _P, _P, _P, _P, _P, _P, _P, _P],
    "km_correct": [_CTX, _GRID, _P, _P, _P, _P, _P, _P, _P, _I],
    "km_mdiv": [_CTX, _GRID, _P, _P, _P, _P],
    "km_time_forward": [_CTX, _GRID, _P, _P, _P, _P],
    "km_stencil_apply": [_CTX, _I, _I, _I, _P, _P, _P],
    "km_dot": [_CTX, ctypes.c_int64, _I, _P, _P, ctypes.POINTER(_D)],
    "km_residual": [_CTX, _I, _I, _P, _P, _P, _P, ctypes.POINTER(_D)],
    "km_bicgstab": [_CTX, _I, _I, _P, _P, _P, _D, _I, _I, _INFO],
    "km_bicgstab_noalloc": [_CTX, _I, _I, _P, _P, _P, _D, _I, _I, _INFO],
    "km_cg": [_CTX, _I, _I, _P, _P, _P, _D, _I, _I, _INFO],
    "km_time_kernel": [_CTX, _I, _I, _I, _P, _P, _I, ctypes.POINTER(_D)],
    "km_unet_param_count": [ctypes.POINTER(ctypes.c_int64)],
    "km_unet_apply": [_CTX, _P, _I, _I, _P, _P, _P],
    "km_unet_loss_grad": [_CTX, _P, _I, _I, _I, _P, _P, _P, _P, ctypes.POINTER(_D)],
    "km_adam_step": [_CTX, ctypes.c_int64, _P, _P, _P, _P, _I, _D, _D, _D, _D],
    "km_fbicgstab": [_CTX, _I, _I, _P, _P, _P, _D, _I, _I, _P, _INFO],
    "km_mg_vcycle": [_CTX, _I, _I, _P, _P, _P, _P],
    "km_fbicgstab2": [_CTX, _I, _I, _P, _P, _P, _I, _I, _P, _P, _P, _D, _I, _I, _INFO, _INFO],
    "km_batch_step": [_CTX, _I, _P, _D, _D, _I, _I, _P],
    "km_simple_step": [_CTX, _P, _D, _D, _I, _I, _P],
    "km_simple_step_n": [_CTX, _I, _P, _D, _D, _P, _P, _P],
    "km_fbicgstab_n": [_CTX, _I, _P, _P, _P, _P, _P, _D, _I, _I, _P],
    "km_batch_time_forward": [_CTX, _I, _P],
    "km_batch_get_tile": [_CTX, ctypes.POINTER(_I), ctypes.POINTER(_I)],
    "km_set_coarse": [_CTX, _I],
    "km_coarse_correct": [_CTX, _I, _I, _P, _P, _P, _P],
    "km_coarse_info": [_CTX, ctypes.POINTER(_I), ctypes.POINTER(_I)],
    "km_set_coarse_update": [_CTX, _I],
    "km_twolevel_precond": [_CTX, _I, _I, _P, _P, _P],
    "km_twolevel_tile": [_CTX, ctypes.POINTER(_I), ctypes.POINTER(_I)],
    "km_twolevel_tile_threads": [_CTX, ctypes.POINTER(_I)],
    "km_twolevel_set_tile": [_CTX, _I],
    "km_set_coarse_totals": [_CTX, _I],
    "km_set_plain_totals": [_CTX, _I],
    "km_set_momentum_sweeps": [_CTX, _I],
    "km_set_coarse_chunks": [_CTX, _I, _I, _I],
}
_RESTYPE = {"km_last_error": ctypes.c_char_p}

_lib = None
_lock = threading.Lock()


class KarmanError(RuntimeError):
    pass


def lib():
    """Load libkarman_hip.so and declare its signatures (raises if it is not built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise KarmanError(
                    f"{LIB_PATH} is not built; run `python -m karman_taichi_amd.build` "
                    "(there is no fallback implementation)")
            h = ctypes.CDLL(str(LIB_PATH))
            for name, args in SIGNATURES.items():
                f = getattr(h, name)
                f.argtypes = args
                f.restype = _RESTYPE.get(name, ctypes.c_int)
            _lib = h
    return _lib


def unet_param_count() -> int:
    """Number of FP32 network parameters (a host-only call)."""
    n = ctypes.c_int64()
    if lib().km_unet_param_count(ctypes.byref(n)) != KM_OK:
        raise KarmanError("km_unet_param_count failed")
    return n.value


def exported_symbols():
    """Names this binding expects the library to export."""
    return list(SIGNATURES)


class Context:
    """One km_ctx bound to a device and a HIP stream (a torch stream handle)."""

    def __init__(self, device: int = 0, stream: int | None = None):
        self._lib = lib()
        h = ctypes.c_void_p()
        rc = self._lib.km_create(ctypes.byref(h), int(device), ctypes.c_void_p(stream or 0))
        if rc != KM_OK:
            raise KarmanError(f"km_create(device={device}) failed with code {rc} (no HIP device?)")
        self.h = h
        self.device = device

    def check(self, rc: int, what: str) -> int:
        if rc < 0:
            msg = self._lib.km_last_error(self.h)
            raise KarmanError(f"{what}: code {rc}: {msg.decode() if msg else ''}")
        return rc

    def call(self, name: str, *args) -> int:
        return self.check(getattr(self._lib, name)(self.h, *args), name)

    def set_stream(self, stream: int):
        self.call("km_set_stream", ctypes.c_void_p(stream or 0))

    def synchronize(self):
        self.call("km_synchronize")

    def close(self):
        if getattr(self, "h", None):
            self._lib.km_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass


_contexts = {}


def context_for(device) -> Context:
    """Process-wide context per device, bound to torch's current stream."""
    import torch
    dev = torch.device(device)
    idx = dev.index if dev.index is not None else torch.cuda.current_device()
    stream = torch.cuda.current_stream(idx).cuda_stream
    ctx = _contexts.get(idx)
    if ctx is None:
        ctx = Context(idx, stream)
        _contexts[idx] = ctx
    else:
        ctx.set_stream(stream)
    return ctx


def ptr(t) -> ctypes.c_void_p:
    """Device pointer of a torch tensor (must be on a HIP device)."""
    if not t.is_cuda:
        raise KarmanError("tensor is not on a HIP device")
    return ctypes.c_void_p(t.data_ptr())
