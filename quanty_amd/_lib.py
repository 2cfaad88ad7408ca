"""ctypes binding of libqs (include/qs.h).

The shared library is built in-tree (quanty_amd/libqs.so) by
``__graft_entry__.build()``. Loading fails loudly when it is missing: there is
no CPU fallback for any operation of this package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqs.so")

QS_OK = 0
QS_HOP_MATFREE, QS_HOP_ASSEMBLED, QS_HOP_AUTO = 0, 1, 2
QS_ENOTCONVERGED = 5
QS_PARAM_SHIFT_MIN, QS_PARAM_ELL_PACK, QS_PARAM_BATCH_GRAPH, QS_PARAM_TERM_LDS, QS_PARAM_ELL_ORDER = 1, 2, 3, 4, 5
QS_PARAM_ALLOC_TOP = 6
QS_PARAM_HETRD = 7
QS_PARAM_TRIDIAG_DC = 8
QS_PARAM_WAIT_SPINS = 9
QS_BATCH_KEEP_BASIS = 1
STATUS = {0: "QS_OK", 1: "QS_ENOMEM_DEVICE", 2: "QS_ENOMEM_HOST", 3: "QS_EINVAL",
          4: "QS_EHIP", 5: "QS_ENOTCONVERGED", 6: "QS_EUNSUPPORTED"}

# (name, restype, argtypes) for every entry point declared in include/qs.h
_P = C.c_void_p
_I64 = C.c_int64
_D = C.c_double
_SIGS = [
    ("qs_ctx_create", C.c_int, [C.c_int, C.POINTER(_P)]),
    ("qs_ctx_destroy", C.c_int, [_P]),
    ("qs_last_error", C.c_char_p, [_P]),
    ("qs_dev_alloc", C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    ("qs_dev_free", C.c_int, [_P, _P]),
    ("qs_trim_memory", C.c_int, [_P, C.c_size_t]),
    ("qs_memory_stats", C.c_int, [_P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    ("qs_memcpy_h2d", C.c_int, [_P, _P, _P, C.c_size_t]),
    ("qs_memcpy_d2h", C.c_int, [_P, _P, _P, C.c_size_t]),
    ("qs_sync", C.c_int, [_P]),
    ("qs_ctx_stream", _P, [_P]),
    ("qs_ctx_set_async", C.c_int, [_P, C.c_int]),
    ("qs_ctx_set_param", C.c_int, [_P, C.c_int, _I64]),
    ("qs_ctx_get_param", C.c_int, [_P, C.c_int, C.POINTER(_I64)]),
    ("qs_basis_enumerate", C.c_int, [_P, C.c_int, C.c_int, _P, _P, _P, C.POINTER(_P)]),
    ("qs_basis_count", C.c_int, [_P, C.c_int, C.c_int, _P, _P, _P, C.POINTER(_I64)]),
    ("qs_basis_from_words", C.c_int, [_P, C.c_int, _P, _I64, C.POINTER(_P)]),
    ("qs_basis_closure", C.c_int, [_P, _P, _P, C.c_int, _P, _P, _P, _I64, C.POINTER(_P)]),
    ("qs_basis_image", C.c_int, [_P, _P, _P, C.c_int, _P, _P, _P, _I64, C.POINTER(_P)]),
    ("qs_basis_size", _I64, [_P]),
    ("qs_basis_nf", C.c_int, [_P]),
    ("qs_basis_words", C.c_int, [_P]),
    ("qs_basis_export", C.c_int, [_P, _P, _P]),
    ("qs_basis_device_words", _P, [_P]),
    ("qs_basis_index", C.c_int, [_P, _P, _P, _I64, _P]),
    ("qs_basis_destroy", C.c_int, [_P]),
    ("qs_op_create", C.c_int, [_P, C.c_int, C.c_int, _P, _P, _P, C.POINTER(_P)]),
    ("qs_op_nterms", C.c_int, [_P]),
    ("qs_op_destroy", C.c_int, [_P]),
    ("qs_hop_create", C.c_int, [_P, _P, _P, _P, C.c_int, C.c_size_t, C.POINTER(_P)]),
    ("qs_hop_create_like", C.c_int, [_P, _P, _P, C.POINTER(_P)]),
    ("qs_hop_create_product", C.c_int, [_P, _P, _P, _P, C.POINTER(_P)]),
    ("qs_hop_row_order", C.c_int, [_P, _P, _P]),
    ("qs_hop_product_info", C.c_int, [_P, _P, C.POINTER(C.c_double)]),
    ("qs_hop_apply_rows", C.c_int, [_P, _P, _P, _P]),
    ("qs_hop_stats", C.c_int, [_P, C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_I64), C.POINTER(C.c_int),
                               C.POINTER(_I64)]),
    ("qs_hop_layout", C.c_int, [_P, C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_I64), C.POINTER(C.c_int),
                                C.POINTER(C.c_int)]),
    ("qs_hop_diagonal", C.c_int, [_P, _P, _P]),
    ("qs_hop_apply", C.c_int, [_P, _P, _P, _P]),
    ("qs_hop_apply_multi", C.c_int, [_P, _P, C.c_int, _P, _P]),
    ("qs_hop_destroy", C.c_int, [_P]),
    ("qs_apply", C.c_int, [_P, _P, _P, _P, _P, _P]),
    ("qs_expect", C.c_int, [_P, _P, _P, _P, C.POINTER(_D), C.POINTER(_D)]),
    ("qs_expect_noalloc", C.c_int, [_P, _P, _P, _P, _P]),
    ("qs_dot", C.c_int, [_P, _I64, _P, _P, C.POINTER(_D), C.POINTER(_D)]),
    ("qs_norm", C.c_int, [_P, _I64, _P, C.POINTER(_D)]),
    ("qs_lanczos", C.c_int, [_P, _P, _P, C.c_int, _P, _P]),
    ("qs_lanczos_krylov", C.c_int, [_P, _P, _P, C.c_int, C.c_double, _P, _P, _P]),
    ("qs_block_lanczos", C.c_int, [_P, _P, C.c_int, _P, C.c_int, _D, _P, _P, _P]),
    ("qs_ground_states", C.c_int, [_P, _P, C.c_int, _D, C.c_int, C.c_int, C.c_uint64, _P, _P, _P, _P, _P]),
    ("qs_ground_states_ex", C.c_int, [_P, _P, C.c_int, _D, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_int,
                                      _P, _P, _P, _P, _P]),
    ("qs_cf_spectrum", C.c_int, [_P, C.c_int, C.c_int, _P, _P, _P, _D, _D, C.c_int, _D, _P, _P]),
    ("qs_block_cf_spectrum", C.c_int, [_P, C.c_int, C.c_int, _P, _P, _P, _D, _D, _D, C.c_int, _D, _P]),
    ("qs_broaden", C.c_int, [_P, C.c_int, C.c_int, _D, C.c_int, C.c_int, _P, _P, _P]),
    ("qs_hop_dense", C.c_int, [_P, _P, _P]),
    ("qs_dense_poles", C.c_int, [_P, C.c_int, _P, C.c_int, _P, _P, _P, C.c_int]),
    ("qs_dense_lowest", C.c_int, [_P, C.c_int, _P, C.c_int, _P, _P]),
    ("qs_pole_spectrum", C.c_int, [_P, C.c_int, C.c_int, _P, _P, C.c_int, _P, _D, _D, C.c_int, _D, _P]),
    ("qs_heev_host", C.c_int, [C.c_int, _P, _P, _P]),
    ("qs_tridiag_lowest_host", C.c_int, [C.c_int, _P, _P, C.POINTER(_D), _P]),
    ("qs_tridiag_lowest_from", C.c_int, [C.c_int, _P, _P, _D, _D, C.POINTER(_D), _P]),
    ("qs_tridiag_extremes_host", C.c_int, [C.c_int, _P, _P, C.POINTER(_D), C.POINTER(_D)]),
    ("qs_batch_create", C.c_int, [_P, C.c_int, C.POINTER(_P)]),
    ("qs_batch_add", C.c_int, [_P, _P, _P, C.c_int, C.c_int, _P, C.POINTER(C.c_int)]),
    ("qs_batch_add_ex", C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.c_int, _P, C.POINTER(C.c_int)]),
    ("qs_batch_ritz", C.c_int, [_P, C.c_int, C.c_int, _P, _P]),
    ("qs_batch_set_depth", C.c_int, [_P, C.c_int, C.c_int]),
    ("qs_batch_run", C.c_int, [_P, C.c_int]),
    ("qs_batch_poll", C.c_int, [_P, C.c_int, C.POINTER(C.c_int), _P, _P, _P]),
    ("qs_batch_remove", C.c_int, [_P, C.c_int]),
    ("qs_batch_active", C.c_int, [_P]),
    ("qs_batch_steps", C.c_int, [_P, _P]),
    ("qs_batch_coeffs", C.c_int, [_P, C.c_int, _P, C.c_int, _P, _P, _P]),
    ("qs_batch_set_depths", C.c_int, [_P, C.c_int, _P, _P]),
    ("qs_batch_destroy", C.c_int, [_P]),
]
EXPORTED = [s[0] for s in _SIGS]

_lib = None


class QSError(RuntimeError):
    pass


def load(path: str = LIB_PATH):
    """Load libqs.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise QSError(f"libqs not built: {path} is missing (run __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, res, args in _SIGS:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def ptr(a) -> int:
    """Raw address of a numpy array or a torch tensor."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


class Context:
    """A libqs context on one device (owns a HIP stream)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        st = self.lib.qs_ctx_create(device, C.byref(h))
        if st != QS_OK:
            raise QSError(f"qs_ctx_create({device}) failed: {STATUS.get(st, st)}")
        self.h = h
        self.device = device

    def check(self, st: int, what: str = ""):
        if st != QS_OK:
            msg = self.lib.qs_last_error(self.h)
            raise QSError(f"{what}: {STATUS.get(st, st)}: {msg.decode() if msg else ''}")

    def set_param(self, param: int, value: int):
        self.check(self.lib.qs_ctx_set_param(self.h, param, int(value)), "qs_ctx_set_param")

    def get_param(self, param: int) -> int:
        v = C.c_int64()
        self.check(self.lib.qs_ctx_get_param(self.h, param, C.byref(v)), "qs_ctx_get_param")
        return v.value

    def memory_stats(self) -> tuple[int, int]:
        """(bytes in use, bytes cached) of the library's device memory on this device."""
        live, cached = C.c_size_t(), C.c_size_t()
        self.check(self.lib.qs_memory_stats(self.h, C.byref(live), C.byref(cached)), "qs_memory_stats")
        return live.value, cached.value

    def trim_memory(self, keep: int = 0):
        """Return cached device memory to the driver until at most `keep` bytes stay cached."""
        self.check(self.lib.qs_trim_memory(self.h, int(keep)), "qs_trim_memory")

    def close(self):
        if getattr(self, "h", None):
            self.lib.qs_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx
