"""ctypes binding of libqtn (the gfx950 C-ABI declared in include/qtn.h).

The library is loaded from this package directory.  There is no fallback:
if libqtn.so is missing or cannot be loaded, every GPU entry point raises
QtnLibraryError.
"""
import ctypes
import os
import threading
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqtn.so")

QTN_OK = 0
QTN_ERR_ARG = -1
QTN_ERR_HIP = -2
QTN_ERR_OOM = -3
QTN_ERR_NOCONV = -4
QTN_ERR_TRUNC = -5
QTN_ERR_NOTFINITE = -6

QTN_FORM_PLAIN = 0
QTN_FORM_CHOSEN = 1
QTN_FORM_LEARNED = 2

# Every symbol include/qtn.h declares (checked by the CPU test suite).
EXPORTS = [
    "qtn_ctx_create", "qtn_ctx_destroy", "qtn_ctx_sync", "qtn_ctx_stream",
    "qtn_ctx_counters", "qtn_ctx_profile", "qtn_ctx_profile_read",
    "qtn_last_error", "qtn_version",
    "qtn_contract_pair", "qtn_contract_pair_dev", "qtn_permute",
    "qtn_contract_network", "qtn_contract_pair_dev_noalloc",
    "qtn_permute_dev", "qtn_zgemm_dev", "qtn_cgemm_dev",
    "qtn_svd_trunc", "qtn_svd", "qtn_svd_values", "qtn_contract_svd",
    "qtn_mps_create", "qtn_mps_destroy", "qtn_mps_copy", "qtn_mps_set_sites",
    "qtn_mps_bond_dims", "qtn_mps_bond_spectrum", "qtn_mps_get_site",
    "qtn_mps_apply_1q", "qtn_mps_apply_2q", "qtn_mps_apply_layer",
    "qtn_mps_switch", "qtn_mps_permute", "qtn_mps_apply_2q_long",
    "qtn_mps_apply_layer_multi", "qtn_mps_apply_1q_multi",
    "qtn_mps_apply_layer_learned", "qtn_mps_gate_features",
    "qtn_gate_schmidt", "qtn_tune_jacobi", "qtn_set_jacobi_tiling",
    "qtn_mps_apply_layer_predicted",
    "qtn_mps_amplitudes", "qtn_mps_expect_z", "qtn_mps_overlap",
    "qtn_mps_norm2", "qtn_mpo_create", "qtn_mpo_destroy",
    "qtn_mps_expect_mpo", "qtn_mps_apply_mpo", "qtn_mps_expect_mpo_form",
    "qtn_mps_apply_mpo_form",
    "qtn_zgemm_bf16x3",
    "qtn_mlp_create", "qtn_mlp_destroy", "qtn_mlp_get_params",
    "qtn_mlp_forward", "qtn_mlp_loss_grad", "qtn_mlp_train",
]


class QtnLibraryError(RuntimeError):
    """libqtn.so is missing or failed to load."""


class QtnError(RuntimeError):
    """A libqtn call returned a non-zero status."""

    def __init__(self, status, message):
        super().__init__("%s (status %d)" % (message, status))
        self.status = status
        self.message = message


_lib = None
_lock = threading.Lock()

_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_vp = ctypes.c_void_p
_pi64 = ctypes.POINTER(ctypes.c_int64)
_pi32 = ctypes.POINTER(ctypes.c_int32)
_pd = ctypes.POINTER(ctypes.c_double)
_pu8 = ctypes.POINTER(ctypes.c_uint8)
_pf = ctypes.POINTER(ctypes.c_float)


def _declare(lib):
    sig = {
        "qtn_ctx_create": (_i32, [_i32, ctypes.POINTER(_vp)]),
        "qtn_ctx_destroy": (_i32, [_vp]),
        "qtn_ctx_sync": (_i32, [_vp]),
        "qtn_ctx_stream": (_i32, [_vp, ctypes.POINTER(_vp)]),
        "qtn_ctx_counters": (_i32, [_vp, _pi64, _i32, _i32]),
        "qtn_ctx_profile": (_i32, [_vp, _i32]),
        "qtn_ctx_profile_read": (_i32, [_vp, _i32, _pi64, _pd, _pd, _pd]),
        "qtn_last_error": (ctypes.c_char_p, []),
        "qtn_version": (ctypes.c_char_p, []),
        "qtn_contract_pair": (_i32, [_vp, _pd, _i32, _pi64, _pd, _i32, _pi64,
                                     _i32, _pi32, _pi32, _pi32, _pd]),
        "qtn_contract_pair_dev": (_i32, [_vp, _vp, _i32, _pi64, _vp, _i32,
                                         _pi64, _i32, _pi32, _pi32, _pi32,
                                         _vp]),
        "qtn_permute": (_i32, [_vp, _pd, _i32, _pi64, _pi32, _pd]),
        "qtn_contract_pair_dev_noalloc": (_i32, [_vp, _vp, _i32, _pi64, _vp,
                                                 _i32, _pi64, _i32, _pi32,
                                                 _pi32, _pi32, _vp]),
        "qtn_contract_network": (_i32, [_vp, _i32, _pi32, _pi64, _pi32,
                                        ctypes.POINTER(_pd), _i32, _pi32,
                                        _pd, _i64]),
        "qtn_permute_dev": (_i32, [_vp, _vp, _i32, _pi64, _pi32, _vp]),
        "qtn_zgemm_dev": (_i32, [_vp, ctypes.c_char, ctypes.c_char, _i64,
                                 _i64, _i64, _pd, _vp, _i64, _i64, _vp, _i64,
                                 _i64, _pd, _vp, _i64, _i64, _i64]),
        "qtn_cgemm_dev": (_i32, [_vp, ctypes.c_char, ctypes.c_char, _i64,
                                 _i64, _i64, ctypes.POINTER(ctypes.c_float),
                                 _vp, _i64, _i64, _vp, _i64, _i64,
                                 ctypes.POINTER(ctypes.c_float), _vp, _i64,
                                 _i64, _i64]),
        "qtn_svd_trunc": (_i32, [_vp, _i64, _i64, _pd, _i64, _dbl, _i32, _i64,
                                 _pi64, _pd, _pd, _pd, _pd]),
        "qtn_svd": (_i32, [_vp, _i64, _i64, _pd, _i64, _i32, _pd, _pd, _pd]),
        "qtn_svd_values": (_i32, [_vp, _i64, _i64, _pd, _i64, _pd]),
        "qtn_contract_svd": (_i32, [_vp, _pd, _i32, _pi64, _pd, _i32, _pi64,
                                    _i32, _i32, _dbl, _pd]),
        "qtn_mps_create": (_i32, [_vp, _i32, _i64, ctypes.POINTER(_vp)]),
        "qtn_mps_destroy": (_i32, [_vp]),
        "qtn_mps_copy": (_i32, [_vp, _vp]),
        "qtn_mps_set_sites": (_i32, [_vp, _pi64, ctypes.POINTER(_pd)]),
        "qtn_mps_bond_dims": (_i32, [_vp, _pi64]),
        "qtn_mps_bond_spectrum": (_i32, [_vp, _i32, _pd, _pi64]),
        "qtn_mps_get_site": (_i32, [_vp, _i32, _pd, _pi64]),
        "qtn_mps_apply_1q": (_i32, [_vp, _i32, _pd]),
        "qtn_mps_apply_2q": (_i32, [_vp, _i32, _pd, _dbl, _i64, _i32, _pi64,
                                    _pd, _pd, _pi64]),
        "qtn_mps_apply_layer": (_i32, [_vp, _i32, _pi32, _pd, _dbl, _i64,
                                       _i32, _pi64, _pd]),
        "qtn_mps_switch": (_i32, [_vp, _i32, _i32]),
        "qtn_mps_permute": (_i32, [_vp, _pi64, _i64]),
        "qtn_mps_apply_2q_long": (_i32, [_vp, _i32, _i32, _pd, _dbl, _i64,
                                         _i32, _pi64, _pd]),
        "qtn_mps_apply_layer_multi": (_i32, [_i32, ctypes.POINTER(_vp), _pi32,
                                             _pd, _pi64, _dbl, _i32, _pi64,
                                             _pd]),
        "qtn_mps_apply_1q_multi": (_i32, [_i32, ctypes.POINTER(_vp), _pi32,
                                          _pd]),
        "qtn_mps_apply_layer_learned": (_i32, [_vp, _i32, _pi32, _pd, _dbl,
                                               _i64, _pi64, _pi64, _pd,
                                               _pi32]),
        "qtn_gate_schmidt": (_i32, [_vp, _i32, _pd, _pd]),
        "qtn_tune_jacobi": (_i32, [_vp, _i64, _i64, _i32, _i32, _pi32, _pd]),
        "qtn_set_jacobi_tiling": (_i32, [_vp, _i64, _i32, _i32, _i32]),
        "qtn_mps_gate_features": (_i32, [_vp, _i32, _pi32, _pd, _i64,
                                         _pf]),
        "qtn_mps_apply_layer_predicted": (_i32, [_vp, _i32, _pi32, _pd, _pd,
                                                 _dbl, _i64, _vp, _dbl,
                                                 _pi64, _pd, _pi32, _pi64]),
        "qtn_mps_amplitudes": (_i32, [_vp, _i64, _pu8, _pd]),
        "qtn_mps_expect_z": (_i32, [_vp, _pd]),
        "qtn_mps_overlap": (_i32, [_vp, _vp, _pd]),
        "qtn_mps_norm2": (_i32, [_vp, _pd]),
        "qtn_mpo_create": (_i32, [_vp, _i32, _pi64, ctypes.POINTER(_pd),
                                  ctypes.POINTER(_vp)]),
        "qtn_mpo_destroy": (_i32, [_vp]),
        "qtn_mps_expect_mpo": (_i32, [_vp, _vp, _pd]),
        "qtn_mps_expect_mpo_form": (_i32, [_vp, _vp, _i32, _pd]),
        "qtn_zgemm_bf16x3": (_i32, [_vp, ctypes.c_char, _i64, _i64, _i64, _pd,
                                    _i64, _pd, _i64, _pd, _i64]),
        "qtn_mps_apply_mpo": (_i32, [_vp, _vp, _dbl, _i64, _vp, _pd]),
        "qtn_mps_apply_mpo_form": (_i32, [_vp, _vp, _dbl, _i64, _i32, _vp,
                                          _pd]),
        "qtn_mlp_create": (_i32, [_vp, _i32, _i32, _pf, ctypes.POINTER(_vp)]),
        "qtn_mlp_destroy": (_i32, [_vp]),
        "qtn_mlp_get_params": (_i32, [_vp, _pf]),
        "qtn_mlp_forward": (_i32, [_vp, _i64, _pf, _pf]),
        "qtn_mlp_loss_grad": (_i32, [_vp, _i64, _pf, _pf, _pf, _pf]),
        "qtn_mlp_train": (_i32, [_vp, _i64, _pf, _pf, _i32, _i64,
                                 ctypes.c_float, _pf]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib_path():
    return LIB_PATH


def load():
    """Load libqtn.so (once).  Raises QtnLibraryError if it is absent."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise QtnLibraryError(
                "libqtn.so not found at %s; build it with "
                "`python qaintensor.jl_amd/build_lib.py` (or "
                "__graft_entry__.build())" % LIB_PATH)
        try:
            lib = ctypes.CDLL(LIB_PATH)
        except OSError as e:
            raise QtnLibraryError("cannot load %s: %s" % (LIB_PATH, e))
        _declare(lib)
        _lib = lib
        return lib


def check(status):
    """Raise QtnError for a non-zero libqtn status."""
    if status != QTN_OK:
        msg = load().qtn_last_error().decode("utf-8", "replace")
        raise QtnError(status, msg)


def dptr(a):
    """Pointer to a C-contiguous numpy array as double*."""
    return a.ctypes.data_as(_pd)


def i64ptr(a):
    return a.ctypes.data_as(_pi64)


def i32ptr(a):
    return a.ctypes.data_as(_pi32)


def as_c128(a, order="F"):
    """complex128 array in the requested memory order (Julia = 'F')."""
    return np.asarray(a, dtype=np.complex128, order=order)


class Context:
    """Owns one qtn_ctx (device, stream, scratch).  The states, operators
    and networks made on it register themselves, and close() frees those
    still alive before the context: the library's objects use their context
    when they are destroyed."""

    def __init__(self, device=0):
        lib = load()
        h = _vp()
        check(lib.qtn_ctx_create(int(device), ctypes.byref(h)))
        self._h = h
        self.device = device
        self._children = weakref.WeakSet()

    def adopt(self, obj):
        """Register obj (with a free() method) to be freed by close()."""
        self._children.add(obj)

    @property
    def handle(self):
        if self._h is None:
            raise QtnError(QTN_ERR_ARG, "context already destroyed")
        return self._h

    def sync(self):
        check(load().qtn_ctx_sync(self.handle))

    def stream(self):
        s = _vp()
        check(load().qtn_ctx_stream(self.handle, ctypes.byref(s)))
        return s.value

    def counters(self, reset=False):
        """dict of the context's work counters (qtn_ctx_counters)."""
        out = np.zeros(6, dtype=np.int64)
        check(load().qtn_ctx_counters(self.handle, i64ptr(out), 6,
                                      int(bool(reset))))
        return dict(svd=int(out[0]), sweeps=int(out[1]),
                    max_sweeps=int(out[2]), batches=int(out[3]),
                    rr_accepted=int(out[4]), rr_redone=int(out[5]))

    PROFILED_KERNELS = ("jac_inner", "jac_gram", "jac_update", "jac_round",
                        "jac_inner_fp32", "jac_gram_fp32",
                        "jac_update_fp32", "jac_round_fp32")

    def profile(self, enable=True):
        """Enable (and clear) or disable per-kernel HIP-event timing."""
        check(load().qtn_ctx_profile(self.handle, int(bool(enable))))

    def profile_read(self):
        """{kernel: dict(count, ms, bytes, flops)} of the profiled kernels."""
        out = {}
        for kid, name in enumerate(self.PROFILED_KERNELS):
            c = ctypes.c_int64(0)
            ms, by, fl = ctypes.c_double(0), ctypes.c_double(0), \
                ctypes.c_double(0)
            check(load().qtn_ctx_profile_read(
                self.handle, kid, ctypes.byref(c), ctypes.byref(ms),
                ctypes.byref(by), ctypes.byref(fl)))
            out[name] = dict(count=c.value, ms=ms.value, bytes=by.value,
                             flops=fl.value)
        return out

    def close(self):
        if self._h is not None:
            for child in list(self._children):
                child.free()
            check(load().qtn_ctx_destroy(self._h))
            self._h = None

    def __del__(self):
        # at interpreter exit contexts may be finalised before their
        # objects: free those first, as close() does
        try:
            if self._h is not None and _lib is not None:
                for child in list(getattr(self, "_children", ())):
                    child.free()
                _lib.qtn_ctx_destroy(self._h)
                self._h = None
        except Exception:
            pass


_default_ctx = None
# held while the default context is created (Context() takes _lock itself
# through load(), so this is a separate lock)
_default_lock = threading.Lock()


def default_context():
    """Process-wide context on device 0 (created on first use; two threads
    asking at once get the same one)."""
    global _default_ctx
    with _default_lock:
        if _default_ctx is None:
            _default_ctx = Context(0)
        return _default_ctx
