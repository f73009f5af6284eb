"""Python front end of libsrhip.so (ctypes over the C ABI in include/sr_hip.h).

Mirrors the reference's evaluate entry points (cl_sr_system.cpp:15-28):
``Context.eval_mse(pop, mode)`` is evaluate_gpu_simple_wrapper (mode PLAIN /
BATCHED), evaluate_gpu_mse_wrapper (mode CACHED) and evaluate_gpu_nvrtc_wrapper
(mode JIT); it returns the per-expression MSE vector and a RunStats-like dict.
Errors raise SrError instead of exiting the process (cl_gpu_common.h:15-23).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

EVAL_PLAIN, EVAL_BATCHED, EVAL_CACHED, EVAL_JIT = 0, 1, 2, 3

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libsrhip.so")

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class SrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sr error {code}: {msg}")
        self.code = code


class _Pop(C.Structure):
    _fields_ = [
        ("num_exprs", C.c_int32), ("tokens", C.c_void_p), ("values", C.c_void_p),
        ("values_f32", C.c_void_p), ("offsets", C.c_void_p), ("lengths", C.c_void_p),
    ]


class _Stats(C.Structure):
    _fields_ = [
        ("total_eval_time_ms", C.c_double), ("detect_time_ms", C.c_double),
        ("transfer_time_ms", C.c_double), ("kernel_time_ms", C.c_double),
        ("jit_compile_time_ms", C.c_double), ("num_subtrees", C.c_int32),
        ("avg_subtree_size", C.c_double), ("coverage", C.c_double),
        ("num_new_subtrees", C.c_int32), ("num_cached", C.c_int32),
        ("num_kernel_launches", C.c_int32), ("logical_node_evals", C.c_int64),
        ("executed_node_evals", C.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class _Tuning(C.Structure):
    _fields_ = [("dpt", C.c_int32), ("chunk", C.c_int32), ("n_queues", C.c_int32),
                ("blocks_per_cu", C.c_int32)]


_lib = None


def lib():
    """Load libsrhip.so; raise if it has not been built (no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SrError(-8, f"{LIB_PATH} is missing: run python -m symbolic_regression_gpu_amd.build")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.sr_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.sr_ctx_destroy.argtypes = [vp]
    L.sr_last_error.argtypes = [vp]
    L.sr_last_error.restype = C.c_char_p
    L.sr_set_stream.argtypes = [vp, vp]
    L.sr_set_data.argtypes = [vp, _f32p, C.c_int32, C.c_int64]
    L.sr_set_data_device.argtypes = [vp, vp, C.c_int32, C.c_int64]
    L.sr_eval_mse.argtypes = [vp, C.POINTER(_Pop), C.c_int, _f32p, _i64p, C.POINTER(_Stats)]
    L.sr_eval_predictions.argtypes = [vp, C.POINTER(_Pop), _f32p]
    L.sr_get_tuning.argtypes = [vp, C.POINTER(_Tuning)]
    L.sr_set_tuning.argtypes = [vp, C.POINTER(_Tuning)]
    L.sr_compile_expr.argtypes = [_i32p, _f32p, C.c_int32, C.c_int32, _i32p, _i32p,
                                  C.POINTER(C.c_int32)]
    _lib = L
    return L


def compile_expr(tokens, values, nv):
    """Bytecode of one expression: (code int32[n], arg int32[n], stack_need)."""
    t = np.ascontiguousarray(tokens, np.int32)
    v = np.ascontiguousarray(values, np.float32)
    code = np.empty(max(len(t), 1), np.int32)
    arg = np.empty(max(len(t), 1), np.int32)
    need = C.c_int32()
    n = lib().sr_compile_expr(t, v, len(t), nv, code, arg, C.byref(need))
    if n < 0:
        raise SrError(n, "compile failed")
    return code[:n], arg[:n], need.value


class Population:
    """Packed prefix-token population (cl_gpu_simple.cu:213-227 layout).

    tokens int32[T], values float64[T] (the doubles InputInfo carries),
    offsets/lengths int32[E].
    """

    def __init__(self, tokens, values, offsets, lengths):
        self.tokens = np.ascontiguousarray(tokens, dtype=np.int32)
        self.values = np.ascontiguousarray(values, dtype=np.float64)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int32)
        self.lengths = np.ascontiguousarray(lengths, dtype=np.int32)

    @classmethod
    def from_exprs(cls, exprs):
        toks, vals, offs, lens = [], [], [], []
        n = 0
        for t, v in exprs:
            offs.append(n)
            lens.append(len(t))
            toks.extend(int(x) for x in t)
            vals.extend(float(x) for x in v)
            n += len(t)
        return cls(toks, vals, offs, lens)

    @classmethod
    def from_strided(cls, tok, val, lengths, cap):
        """From a fixed-stride population (individual i at i*cap)."""
        lengths = np.asarray(lengths, dtype=np.int32)
        idx = np.concatenate([np.arange(i * cap, i * cap + l) for i, l in enumerate(lengths)]) \
            if len(lengths) else np.zeros(0, np.int64)
        offs = np.zeros(len(lengths), np.int32)
        if len(lengths) > 1:
            offs[1:] = np.cumsum(lengths[:-1])
        return cls(np.asarray(tok)[idx], np.asarray(val, dtype=np.float64)[idx], offs, lengths)

    @property
    def num_exprs(self):
        return int(self.lengths.shape[0])

    def expr(self, e):
        a = int(self.offsets[e])
        b = a + int(self.lengths[e])
        return self.tokens[a:b], self.values[a:b]

    def _c(self):
        p = _Pop()
        p.num_exprs = self.num_exprs
        p.tokens = self.tokens.ctypes.data
        p.values = self.values.ctypes.data
        p.values_f32 = None
        p.offsets = self.offsets.ctypes.data
        p.lengths = self.lengths.ctypes.data
        return p


@dataclass
class Tuning:
    dpt: int = 4
    chunk: int = 16384
    n_queues: int = 8
    blocks_per_cu: int = 0


class Context:
    """One evaluator context on one GPU (create_gpu_context analogue)."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = C.c_void_p()
        rc = self._L.sr_ctx_create(device, C.byref(h))
        if rc != 0:
            raise SrError(rc, "sr_ctx_create failed (no HIP device?)")
        self._h = h
        self.nv = None
        self.N = None
        self._keep = None

    def close(self):
        if getattr(self, "_h", None):
            self._L.sr_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise SrError(rc, self._L.sr_last_error(self._h).decode())

    def set_stream(self, stream_ptr: int | None):
        self._check(self._L.sr_set_stream(self._h, stream_ptr))

    def set_data(self, X, y=None):
        """X: [nv, N] (column-major per variable) or [nv+1, N] with y last."""
        X = np.asarray(X, dtype=np.float32)
        if y is not None:
            X = np.vstack([X, np.asarray(y, np.float32)[None, :]])
        X = np.ascontiguousarray(X)
        nv = X.shape[0] - 1
        self._check(self._L.sr_set_data(self._h, X.ravel(), nv, X.shape[1]))
        self.nv, self.N = nv, X.shape[1]

    def set_data_device(self, ptr: int, nv: int, N: int, keepalive=None):
        self._check(self._L.sr_set_data_device(self._h, ptr, nv, N))
        self.nv, self.N, self._keep = nv, N, keepalive

    def eval_mse(self, pop: Population, mode: int = EVAL_BATCHED):
        E = pop.num_exprs
        mse = np.empty(max(E, 1), np.float32)
        valid = np.empty(max(E, 1), np.int64)
        st = _Stats()
        self._check(self._L.sr_eval_mse(self._h, C.byref(pop._c()), mode, mse, valid, C.byref(st)))
        return mse[:E], valid[:E], st.as_dict()

    def eval_predictions(self, pop: Population):
        out = np.empty(max(pop.num_exprs, 1) * self.N, np.float32)
        self._check(self._L.sr_eval_predictions(self._h, C.byref(pop._c()), out))
        return out[:pop.num_exprs * self.N].reshape(pop.num_exprs, self.N)

    @property
    def tuning(self) -> Tuning:
        t = _Tuning()
        self._check(self._L.sr_get_tuning(self._h, C.byref(t)))
        return Tuning(t.dpt, t.chunk, t.n_queues, t.blocks_per_cu)

    @tuning.setter
    def tuning(self, v: Tuning):
        t = _Tuning(v.dpt, v.chunk, v.n_queues, v.blocks_per_cu)
        self._check(self._L.sr_set_tuning(self._h, C.byref(t)))
