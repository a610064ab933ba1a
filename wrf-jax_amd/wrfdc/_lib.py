"""ctypes binding of ``libwrfdc.so`` (the HIP dycore, C ABI in include/wrfdc.h).

The library is built in-tree by ``__graft_entry__.build()`` (hipcc,
``--offload-arch=gfx950``).  There is no fallback: if the shared object is
missing or does not load, every entry point of this package raises.
"""

from __future__ import annotations

import ctypes
import os

from .fields import FIELD_SPECS

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libwrfdc.so")

WRFDC_OK = 0
WRFDC_EINVAL = -1
WRFDC_ENOTIMPL = -2
WRFDC_EUNBOUND = -3
WRFDC_EHIP = -4
WRFDC_ENONFINITE = -5

# every symbol include/wrfdc.h declares, with its ctypes signature
SIGNATURES = {
    "wrfdc_field_count": (ctypes.c_int, []),
    "wrfdc_field_name": (ctypes.c_char_p, [ctypes.c_int]),
    "wrfdc_field_kind": (ctypes.c_int, [ctypes.c_int]),
    "wrfdc_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "wrfdc_destroy": (None, [ctypes.c_void_p]),
    "wrfdc_bind": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]),
    "wrfdc_set_stream": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "wrfdc_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]),
    "wrfdc_stage": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]),
    "wrfdc_set_dtbc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "wrfdc_set_tile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]),
    "wrfdc_checksums": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "wrfdc_checksums_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p,
                                                ctypes.c_void_p]),
    "wrfdc_ensemble_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_void_p)]),
    "wrfdc_ensemble_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]),
    "wrfdc_ensemble_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "wrfdc_ensemble_stream": (ctypes.c_void_p, [ctypes.c_void_p]),
    "wrfdc_ensemble_nodes": (ctypes.c_longlong, [ctypes.c_void_p]),
    "wrfdc_ensemble_destroy": (None, [ctypes.c_void_p]),
    "wrfdc_record_stages": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "wrfdc_recorded_stages": (ctypes.c_int, [ctypes.c_void_p]),
    "wrfdc_transport_stage": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p]),
    "wrfdc_transport_stage_adjoint": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                     ctypes.c_void_p, ctypes.c_void_p]),
    "wrfdc_set_closure": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "wrfdc_closure_row_count": (ctypes.c_longlong, [ctypes.c_void_p]),
    "wrfdc_closure_rows": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "wrfdc_subgrid_stress": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "wrfdc_coarse_grain": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "wrfdc_mlp_work_floats": (ctypes.c_longlong, [ctypes.c_longlong]),
    "wrfdc_mlp_pack": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "wrfdc_mlp_grad": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong,
                                      ctypes.c_float, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p]),
    "wrfdc_mlp_train_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong,
                                            ctypes.c_float, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_int, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p]),
    "wrfdc_mlp_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p,
                                         ctypes.c_void_p]),
    "wrfdc_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "wrfdc_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
}

_LIB: ctypes.CDLL | None = None


class LibraryMissingError(RuntimeError):
    """libwrfdc.so is not built or does not load (no CPU fallback exists)."""


def load() -> ctypes.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise LibraryMissingError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    try:
        lib = ctypes.CDLL(LIB_PATH)
    except OSError as exc:  # pragma: no cover - depends on the ROCm runtime
        raise LibraryMissingError(f"cannot load {LIB_PATH}: {exc}") from exc
    for name, (restype, argtypes) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    table = tuple((lib.wrfdc_field_name(i).decode(), chr(lib.wrfdc_field_kind(i))) for i in range(lib.wrfdc_field_count()))
    if table != FIELD_SPECS:
        raise LibraryMissingError("libwrfdc.so field table differs from wrfdc.fields.FIELD_SPECS (stale build?)")
    _LIB = lib
    return lib


def raise_for(rc: int, lib: ctypes.CDLL, ctx, what: str) -> None:
    """Map a negative return code onto the reference's exception types."""
    if rc == WRFDC_OK:
        return
    raw = lib.wrfdc_last_error(ctx)  # ctx None: the failure of a context-free call on this thread
    msg = raw.decode() if raw else ""
    text = f"{what}: {msg}" if msg else what
    if rc == WRFDC_EINVAL:
        raise ValueError(text)
    if rc == WRFDC_ENOTIMPL:
        raise NotImplementedError(text)
    if rc == WRFDC_EUNBOUND:
        raise KeyError(text)
    if rc == WRFDC_ENONFINITE:
        raise FloatingPointError(text)
    raise RuntimeError(f"{text} (HIP error, code {rc})")


__all__ = ["LIB_PATH", "LibraryMissingError", "SIGNATURES", "load", "raise_for"]
