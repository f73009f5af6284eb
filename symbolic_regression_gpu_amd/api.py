"""Python front end of libsrhip.so (ctypes over the C ABI in include/sr_hip.h).

Mirrors the reference's evaluate entry points (cl_sr_system.cpp:15-28):
``Context.eval_mse(pop, mode)`` is evaluate_gpu_simple_wrapper (mode PLAIN /
BATCHED), evaluate_gpu_mse_wrapper (mode CACHED) and evaluate_gpu_nvrtc_wrapper
(mode JIT); it returns the per-expression MSE vector, the finite-prediction
counts and a RunStats-like dict.  The GA methods replace the host GA of
utils.cpp:1156-1272 and the selection loop of cl_sr_system.cpp:285-330.
Errors raise SrError instead of exiting the process (cl_gpu_common.h:15-23).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

EVAL_PLAIN, EVAL_BATCHED, EVAL_CACHED, EVAL_JIT = 0, 1, 2, 3
CACHE_REFERENCE, CACHE_GENERATIONAL, CACHE_GENERATIONAL_REUSED, CACHE_GENERATIONAL_SHARED = 0, 1, 2, 3
# generational cache planned on the device (GA population: ga_run / ga_evaluate)
CACHE_DEVICE = 4

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libsrhip.so")

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_vp = C.c_void_p


class SrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sr error {code}: {msg}")
        self.code = code


class _Pop(C.Structure):
    _fields_ = [
        ("num_exprs", C.c_int32), ("tokens", _vp), ("values", _vp),
        ("values_f32", _vp), ("offsets", _vp), ("lengths", _vp),
    ]


class _Stats(C.Structure):
    _fields_ = [
        ("total_eval_time_ms", C.c_double), ("detect_time_ms", C.c_double),
        ("transfer_time_ms", C.c_double), ("kernel_time_ms", C.c_double),
        ("jit_compile_time_ms", C.c_double), ("num_subtrees", C.c_int32),
        ("avg_subtree_size", C.c_double), ("coverage", C.c_double),
        ("num_new_subtrees", C.c_int32), ("num_cached", C.c_int32),
        ("num_kernel_launches", C.c_int32), ("logical_node_evals", C.c_int64),
        ("executed_node_evals", C.c_int64), ("interp_kernel_ms", C.c_double),
        ("interp_launches", C.c_int32), ("executed_op_evals", C.c_int64),
        ("column_reads", C.c_int64), ("column_writes", C.c_int64), ("graph_captures", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class _Tuning(C.Structure):
    _fields_ = [("dpt", C.c_int32), ("chunk", C.c_int32), ("n_queues", C.c_int32),
                ("blocks_per_cu", C.c_int32), ("cache_dpt", C.c_int32), ("kernel", C.c_int32),
                ("cache_kernel", C.c_int32)]


class GaParams(C.Structure):
    """sr_ga_params: generate / mutate / crossover settings (utils.h:178-187)."""
    _fields_ = [
        ("num_vars", C.c_int32), ("max_tokens", C.c_int32), ("max_depth", C.c_int32),
        ("const_min", C.c_float), ("const_max", C.c_float),
        ("prob_leaf", C.c_float), ("prob_const_leaf", C.c_float), ("prob_unary", C.c_float),
        ("n_unary", C.c_int32), ("n_binary", C.c_int32),
        ("unary_ops", C.c_int32 * 32), ("binary_ops", C.c_int32 * 32),
    ]

    @classmethod
    def make(cls, **kw):
        p = cls()
        lib().sr_ga_params_default(C.byref(p))
        for k, v in kw.items():
            if k in ("unary_ops", "binary_ops"):
                arr = getattr(p, k)
                for i, o in enumerate(v):
                    arr[i] = int(o)
                setattr(p, "n_unary" if k == "unary_ops" else "n_binary", len(v))
            else:
                setattr(p, k, v)
        return p


class SelectParams(C.Structure):
    """sr_select_params: tournament size, mutation share, elite fraction."""
    _fields_ = [("tournament", C.c_int32), ("p_mutation", C.c_float), ("elite_frac", C.c_double)]


_SIGNATURES = {
    "sr_ctx_create": [C.c_int, C.POINTER(_vp)],
    "sr_ctx_destroy": [_vp],
    "sr_last_error": [_vp],
    "sr_set_stream": [_vp, _vp],
    "sr_use_own_stream": [_vp],
    "sr_set_data": [_vp, _f32p, C.c_int32, C.c_int64],
    "sr_set_data_device": [_vp, _vp, C.c_int32, C.c_int64],
    "sr_eval_mse": [_vp, C.POINTER(_Pop), C.c_int, _f32p, _i64p, C.POINTER(_Stats)],
    "sr_eval_mse_noalloc": [_vp, C.c_int32, C.c_int64, _vp, _vp, _vp, _vp, _vp, _vp],
    "sr_eval_predictions": [_vp, C.POINTER(_Pop), _f32p],
    "sr_get_tuning": [_vp, C.POINTER(_Tuning)],
    "sr_autotune": [_vp, C.POINTER(_Pop), C.c_int32, C.POINTER(_Tuning), C.POINTER(C.c_double)],
    "sr_autotune_cache": [_vp, C.POINTER(_Tuning), C.POINTER(C.c_double)],
    "sr_tune_save": [C.c_char_p],
    "sr_runtime_info": [C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
    "sr_tune_load": [C.c_char_p, C.POINTER(C.c_int32)],
    "sr_param_autotune": [_vp, _vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
    "sr_set_tuning": [_vp, C.POINTER(_Tuning)],
    "sr_compile_expr": [_i32p, _f32p, C.c_int32, C.c_int32, _i32p, _i32p, C.POINTER(C.c_int32)],
    "sr_compile_skeleton": [_i32p, _f32p, C.c_int32, C.c_int32, C.c_int32, _i32p, _i32p, C.POINTER(C.c_int32)],
    "sr_compile_param_expr": [_i32p, _f32p, _i32p, C.c_int32, C.c_int32, _i32p, _i32p, _vp,
                              C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
    "sr_jit_source": [C.POINTER(_Pop), C.c_int32, C.c_int32, _vp, C.c_int64, C.POINTER(C.c_int64)],
    "sr_jit_compile": [C.POINTER(_Pop), C.c_int32, C.c_int32, _vp, C.c_int64,
                       C.POINTER(C.c_double), C.POINTER(C.c_int64)],
    "sr_jit_code": [C.POINTER(_Pop), C.c_int32, C.c_int32, _vp, C.c_int64, C.POINTER(C.c_int64)],
    "sr_cache_config": [_vp, C.c_int32, C.c_int32, C.c_int32],
    "sr_cache_clear": [_vp],
    "sr_cache_policy": [_vp, C.c_int32, C.c_int32, C.c_int32],
    "sr_digest_read": [C.c_char_p, C.POINTER(_vp)],
    "sr_digest_free": [_vp],
    "sr_digest_shape": [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int32)],
    "sr_digest_population": [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "sr_digest_data_path": [_vp, C.c_int32, _vp, C.c_int64, C.POINTER(C.c_int64)],
    "sr_digest_write": [C.c_char_p, C.POINTER(_Pop), _vp, _vp, _vp],
    "sr_data_file_read": [C.c_char_p, C.c_int32, C.c_int64, _vp, C.POINTER(C.c_int64)],
    "sr_data_file_write": [C.c_char_p, C.c_int32, C.c_int64, _vp],
    "sr_evogp_to_prefix": [_vp, C.c_int32, C.c_int32, _vp, _vp, _vp, C.c_int32, _vp, _vp, _vp],
    "sr_prefix_to_evogp": [_vp, C.c_int32, C.c_int32, _vp, _vp, _vp, C.c_int32, _vp, _vp, _vp],
    "sr_ga_import_evogp": [_vp, C.c_int32, C.c_int32, _vp, _vp, _vp],
    "sr_ga_export_evogp": [_vp, C.c_int32, _vp, _vp, _vp],
    "sr_memo_planner_create": [C.c_int32, C.c_int32, C.c_int32, C.POINTER(_vp)],
    "sr_memo_planner_destroy": [_vp],
    "sr_memo_plan": [_vp, C.POINTER(_Pop), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                     C.POINTER(C.c_int64)],
    "sr_memo_plan_copy": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "sr_memo_commit": [_vp, _vp, _vp],
    "sr_memo_remembered": [_vp, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_int64)],
    "sr_cache_state": [_vp, _vp, C.c_int32, C.POINTER(C.c_int32)],
    "sr_cache_last_hints": [_vp, _vp, C.c_int64, C.POINTER(C.c_int64)],
    "sr_detect": [_vp, C.POINTER(_Pop), C.c_int32, C.c_int32, _i32p, _u64p, _i32p, _i32p,
                  C.c_int32, C.POINTER(C.c_int32)],
    "sr_ga_params_default": [C.POINTER(GaParams)],
    "sr_ga_init": [_vp, C.POINTER(GaParams), C.c_int32, C.c_uint64],
    "sr_ga_get_population": [_vp, _vp, _vp, _vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
    "sr_ga_set_population": [_vp, _i32p, _f32p, _i32p, C.c_int32, C.c_int32],
    "sr_ga_evaluate": [_vp, C.c_int, _f32p, C.POINTER(_Stats)],
    "sr_ga_next": [_vp, C.POINTER(SelectParams), _vp, C.c_uint64, C.c_uint32],
    "sr_ga_run": [_vp, C.POINTER(SelectParams), C.c_int, C.c_int32, C.c_uint32, C.c_uint64,
                  C.c_int, _vp, C.POINTER(C.c_double), C.POINTER(_Stats)],
    "sr_param_prepare": [_vp, C.POINTER(_Pop), _i32p, _vp],
    "sr_param_info": [_vp, C.POINTER(C.c_int32), _vp, C.POINTER(C.c_int64)],
    "sr_param_eval": [_vp, _vp, _vp],
    "sr_param_grad": [_vp, _vp, _vp, _vp],
    "sr_grad_consts": [_vp, C.POINTER(_Pop), _vp, _vp, _vp],
    "sr_evolve_step": [_vp, C.POINTER(_Pop), _vp, C.POINTER(GaParams), C.POINTER(SelectParams), C.c_uint64,
                       C.c_uint32, C.c_int32, _vp, _vp, _vp],
    "sr_param_fd_grad": [_vp, _vp, C.c_float, _vp],
    "sr_sync": [_vp],
    "sr_tree_splice": [_vp, C.c_int32, C.c_int32, _i32p, _f32p, _i32p, C.c_int32, _i32p, _i32p,
                       _i32p, _f32p, _i32p, C.c_int32, _i32p, _i32p, _i32p, _f32p, _i32p],
}

_lib = None


def lib():
    """Load libsrhip.so; raise if it has not been built (no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SrError(-8, f"{LIB_PATH} is missing: run python -m symbolic_regression_gpu_amd.build")
    # One HIP runtime per process: torch ships its own libamdhip64.so.7 /
    # libhiprtc.so.7, and whichever copy is loaded first serves every later
    # user of that soname.  Loading torch first makes this library share the
    # runtime that owns torch's tensors and streams (the autograd shim passes
    # both across), whatever the caller imported first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    for name, args in _SIGNATURES.items():
        f = getattr(L, name, None)
        if f is None:
            continue
        f.argtypes = args
    L.sr_last_error.restype = C.c_char_p
    _lib = L
    return L


def compile_expr(tokens, values, nv, const_slot=None):
    """Bytecode of one expression: (code int32[n], arg int32[n], stack_need).
    With const_slot (per token -1 or 0..7), the parametric program the adjoint
    pass runs: (code, arg, stack_need, aux int32[n], tape_rows)."""
    t = np.ascontiguousarray(tokens, np.int32)
    v = np.ascontiguousarray(values, np.float32)
    code = np.empty(max(len(t), 1), np.int32)
    arg = np.empty(max(len(t), 1), np.int32)
    need = C.c_int32()
    if const_slot is None:
        n = lib().sr_compile_expr(t, v, len(t), nv, code, arg, C.byref(need))
        if n < 0:
            raise SrError(n, "compile failed")
        return code[:n], arg[:n], need.value
    sl = np.ascontiguousarray(const_slot, np.int32)
    aux = np.empty(max(len(t), 1), np.int32)
    rows = C.c_int32()
    n = lib().sr_compile_param_expr(t, v, sl, len(t), nv, code, arg, aux.ctypes.data, C.byref(need),
                                    C.byref(rows))
    if n < 0:
        raise SrError(n, "compile failed")
    return code[:n], arg[:n], need.value, aux[:n], rows.value


def compile_skeleton(tokens, values, nv, algorithm=1):
    """Bytecode of one cache skeleton (OP_REF / OP_STORE tokens allowed):
    (code int32[n], arg int32[n], stack_need).  algorithm 0 = the serial
    compiler, 1 = the wave-parallel compiler of the device-planned cache (run
    on the host)."""
    t = np.ascontiguousarray(tokens, np.int32)
    v = np.ascontiguousarray(values, np.float32)
    code = np.empty(max(len(t), 1), np.int32)
    arg = np.empty(max(len(t), 1), np.int32)
    need = C.c_int32()
    n = lib().sr_compile_skeleton(t, v, len(t), nv, algorithm, code, arg, C.byref(need))
    if n < 0:
        raise SrError(n, "compile failed")
    return code[:n], arg[:n], need.value


def jit_source(pop, nv: int, dpt: int = 4) -> str:
    """HIP source the hipRTC backend generates for a population."""
    n = C.c_int64()
    rc = lib().sr_jit_source(C.byref(pop._c()), nv, dpt, None, 0, C.byref(n))
    if rc != 0:
        raise SrError(rc, "jit source")
    buf = C.create_string_buffer(n.value + 1)
    lib().sr_jit_source(C.byref(pop._c()), nv, dpt, buf, n.value + 1, C.byref(n))
    return buf.value.decode()


def runtime_info():
    """(HIP version built against, runtime version loaded, driver version, same major)."""
    b, r, d = C.c_int32(), C.c_int32(), C.c_int32()
    rc = lib().sr_runtime_info(C.byref(b), C.byref(r), C.byref(d))
    return b.value, r.value, d.value, rc == 0


def tune_save(path: str) -> None:
    """Write the remembered tunings (every context of this process) to a text file."""
    rc = lib().sr_tune_save(path.encode())
    if rc != 0:
        raise SrError(rc, f"cannot write tunings to {path}")


def tune_load(path: str) -> int:
    """Add the tunings of a file written by tune_save; returns the entry count."""
    n = C.c_int32(0)
    rc = lib().sr_tune_load(path.encode(), C.byref(n))
    if rc != 0:
        raise SrError(rc, f"cannot read tunings from {path}")
    return n.value


def jit_compile(pop, nv: int, dpt: int = 4):
    """Compile the generated module for gfx950 on the host: (ms, code bytes)."""
    log = C.create_string_buffer(1 << 16)
    ms = C.c_double()
    nb = C.c_int64()
    rc = lib().sr_jit_compile(C.byref(pop._c()), nv, dpt, log, 1 << 16, C.byref(ms), C.byref(nb))
    if rc != 0:
        raise SrError(rc, log.value.decode())
    return ms.value, nb.value


def jit_code(pop, nv: int, dpt: int = 4) -> bytes:
    """The gfx950 code object the hipRTC backend builds for a population."""
    n = C.c_int64()
    rc = lib().sr_jit_code(C.byref(pop._c()), nv, dpt, None, 0, C.byref(n))
    if rc != 0:
        raise SrError(rc, "jit code")
    buf = C.create_string_buffer(n.value)
    rc = lib().sr_jit_code(C.byref(pop._c()), nv, dpt, buf, n.value, C.byref(n))
    if rc != 0:
        raise SrError(rc, "jit code")
    return buf.raw[:n.value]


@dataclass
class Digest:
    """A digest file (gen_k.txt) as parse_input_info reads it."""
    population: "Population"
    num_vars: np.ndarray
    num_dps: np.ndarray
    data_paths: list
    shared: bool


def read_digest(path: str) -> Digest:
    """Parse a digest file (utils.cpp:122-233 format); raises SrError."""
    L = lib()
    h = _vp()
    rc = L.sr_digest_read(os.fsencode(path), C.byref(h))
    if rc != 0:
        raise SrError(rc, f"sr_digest_read({path})")
    try:
        E, T, sh = C.c_int32(), C.c_int64(), C.c_int32()
        L.sr_digest_shape(h, C.byref(E), C.byref(T), C.byref(sh))
        tok = np.empty(T.value, np.int32)
        val = np.empty(T.value, np.float64)
        off = np.empty(E.value, np.int32)
        ln = np.empty(E.value, np.int32)
        nv = np.empty(E.value, np.int32)
        nd = np.empty(E.value, np.int32)
        L.sr_digest_population(h, tok.ctypes.data, val.ctypes.data, off.ctypes.data, ln.ctypes.data,
                               nv.ctypes.data, nd.ctypes.data)
        paths = []
        for e in range(E.value):
            n = C.c_int64()
            L.sr_digest_data_path(h, e, None, 0, C.byref(n))
            buf = C.create_string_buffer(n.value + 1)
            L.sr_digest_data_path(h, e, buf, n.value + 1, C.byref(n))
            paths.append(buf.value.decode())
        return Digest(Population(tok, val, off, ln), nv, nd, paths, bool(sh.value))
    finally:
        L.sr_digest_free(h)


def write_digest(path: str, pop: "Population", num_vars, num_dps, data_paths):
    """Write pop in the digest format (values with 17 significant digits)."""
    E = pop.num_exprs
    nv = np.ascontiguousarray(np.broadcast_to(num_vars, (E,)), np.int32)
    nd = np.ascontiguousarray(np.broadcast_to(num_dps, (E,)), np.int32)
    paths = [data_paths] * E if isinstance(data_paths, str) else list(data_paths)
    arr = (C.c_char_p * E)(*[p.encode() for p in paths])
    rc = lib().sr_digest_write(os.fsencode(path), C.byref(pop._c()), nv.ctypes.data, nd.ctypes.data,
                               C.cast(arr, _vp))
    if rc != 0:
        raise SrError(rc, f"sr_digest_write({path})")


def read_data_file(path: str, num_vars: int, num_dps: int):
    """A data file as load_data_file reads it (utils.cpp:235-300):
    (float64 [num_vars+1, num_dps] column-major X|y, numbers converted)."""
    out = np.empty((num_vars + 1) * num_dps, np.float64)
    n = C.c_int64()
    rc = lib().sr_data_file_read(os.fsencode(path), num_vars, num_dps, out.ctypes.data, C.byref(n))
    if rc != 0:
        raise SrError(rc, f"sr_data_file_read({path})")
    return out.reshape(num_vars + 1, num_dps), n.value


def write_data_file(path: str, cols):
    """Write X|y columns [nv+1, N] as N rows of nv+1 numbers."""
    c = np.ascontiguousarray(cols, np.float64)
    rc = lib().sr_data_file_write(os.fsencode(path), c.shape[0] - 1, c.shape[1], c.ctypes.data)
    if rc != 0:
        raise SrError(rc, f"sr_data_file_write({path})")


class MemoPlanner:
    """Host planner of the generational subtree cache (no device needed):
    turns populations into launches of skeleton programs whose OP_REF (-2)
    tokens read, and OP_STORE (-3) tokens write, value columns by slot."""

    def __init__(self, capacity: int, min_size: int = 4, policy: int = CACHE_GENERATIONAL):
        self._L = lib()
        self._h = _vp()
        rc = self._L.sr_memo_planner_create(capacity, min_size, policy, C.byref(self._h))
        if rc != 0:
            raise SrError(rc, "sr_memo_planner_create")
        self._nprog = 0

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.sr_memo_planner_destroy(self._h)
            self._h = None

    def plan(self, pop: "Population") -> dict:
        nl, npg, nt = C.c_int32(), C.c_int32(), C.c_int64()
        rc = self._L.sr_memo_plan(self._h, C.byref(pop._c()), C.byref(nl), C.byref(npg), C.byref(nt))
        if rc != 0:
            raise SrError(rc, "sr_memo_plan")
        out = {"tokens": np.empty(nt.value, np.int32), "values": np.empty(nt.value, np.float64),
               "offsets": np.empty(npg.value, np.int32), "lengths": np.empty(npg.value, np.int32),
               "launch_begin": np.empty(nl.value + 1, np.int32),
               "root_slot": np.empty(npg.value, np.int32),
               "src": np.empty(pop.num_exprs, np.int32)}
        self._L.sr_memo_plan_copy(self._h, *(out[k].ctypes.data for k in
                                             ("tokens", "values", "offsets", "lengths", "launch_begin",
                                              "root_slot", "src")))
        self._nprog = npg.value
        return out

    def commit(self, mse, cnt):
        m = np.ascontiguousarray(mse, np.float32)
        k = np.ascontiguousarray(cnt, np.int64)
        assert len(m) == len(k) == self._nprog
        rc = self._L.sr_memo_commit(self._h, m.ctypes.data, k.ctypes.data)
        if rc != 0:
            raise SrError(rc, "sr_memo_commit")

    def remembered(self, slot: int):
        m, k = C.c_float(), C.c_int64()
        rc = self._L.sr_memo_remembered(self._h, slot, C.byref(m), C.byref(k))
        if rc != 0:
            raise SrError(rc, "sr_memo_remembered")
        return m.value, k.value


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
        idx = (np.concatenate([np.arange(i * cap, i * cap + n) for i, n in enumerate(lengths)])
               if len(lengths) else np.zeros(0, np.int64))
        offs = np.zeros(len(lengths), np.int32)
        if len(lengths) > 1:
            offs[1:] = np.cumsum(lengths[:-1])
        return cls(np.asarray(tok)[idx], np.asarray(val, dtype=np.float64)[idx], offs, lengths)

    @property
    def num_exprs(self):
        return int(self.lengths.shape[0])

    @property
    def total_tokens(self):
        return int(self.lengths.sum())

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
    dpt: int = 3
    chunk: int = 3072
    n_queues: int = 8
    blocks_per_cu: int = 0
    cache_dpt: int = 3
    kernel: int = 1        # whole expressions: 0 lane interpreter, 1 vector interpreter
    cache_kernel: int = 1  # cache skeletons: 0 lane interpreter, 1 vector interpreter


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

    # ---- data -----------------------------------------------------------
    def set_stream(self, stream_ptr: int | None):
        """Run on a caller's hipStream_t handle (0 / None = the null stream,
        torch's default stream)."""
        self._check(self._L.sr_set_stream(self._h, stream_ptr))

    def use_own_stream(self):
        self._check(self._L.sr_use_own_stream(self._h))

    def set_data(self, X, y=None):
        """X: [nv, N] (one row per variable) or [nv+1, N] with y last."""
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

    # ---- evaluation -----------------------------------------------------
    def eval_mse(self, pop: Population, mode: int = EVAL_BATCHED):
        E = pop.num_exprs
        mse = np.empty(max(E, 1), np.float32)
        valid = np.empty(max(E, 1), np.int64)
        st = _Stats()
        self._check(self._L.sr_eval_mse(self._h, C.byref(pop._c()), mode, mse, valid, C.byref(st)))
        return mse[:E], valid[:E], st.as_dict()

    def eval_mse_noalloc(self, tok, val, off, ln, mse, valid=None):
        """sr_eval_mse_noalloc on device tensors (int32 tok / off / ln, float32
        val, float32 mse [E], int64 valid [E] or None); only enqueues on the
        context stream -- call sync() before reading the outputs."""
        self._check(self._L.sr_eval_mse_noalloc(
            self._h, int(ln.numel()), int(tok.numel()), tok.data_ptr(), val.data_ptr(), off.data_ptr(),
            ln.data_ptr(), mse.data_ptr(), valid.data_ptr() if valid is not None else None))

    def eval_predictions(self, pop: Population):
        out = np.empty(max(pop.num_exprs, 1) * self.N, np.float32)
        self._check(self._L.sr_eval_predictions(self._h, C.byref(pop._c()), out))
        return out[:pop.num_exprs * self.N].reshape(pop.num_exprs, self.N)

    @property
    def tuning(self) -> Tuning:
        t = _Tuning()
        self._check(self._L.sr_get_tuning(self._h, C.byref(t)))
        return Tuning(t.dpt, t.chunk, t.n_queues, t.blocks_per_cu, t.cache_dpt, t.kernel, t.cache_kernel)

    @tuning.setter
    def tuning(self, v: Tuning):
        t = _Tuning(v.dpt, v.chunk, v.n_queues, v.blocks_per_cu, v.cache_dpt, v.kernel, v.cache_kernel)
        self._check(self._L.sr_set_tuning(self._h, C.byref(t)))

    # ---- subtree cache ----------------------------------------------------
    def autotune(self, pop: Population, max_sample: int = 256):
        """Time the launch shapes of the batched interpreter on a sample of pop;
        keeps the fastest as this context's tuning.  Returns (Tuning, ms)."""
        t = _Tuning()
        ms = C.c_double()
        self._check(self._L.sr_autotune(self._h, C.byref(pop._c()), max_sample, C.byref(t), C.byref(ms)))
        return Tuning(t.dpt, t.chunk, t.n_queues, t.blocks_per_cu, t.cache_dpt, t.kernel, t.cache_kernel), ms.value

    def autotune_cache(self):
        """Time the skeleton kernels of the device-planned cache on the last
        cached generation's plan; keeps the fastest.  Returns (Tuning, ms)."""
        t = _Tuning()
        ms = C.c_double()
        self._check(self._L.sr_autotune_cache(self._h, C.byref(t), C.byref(ms)))
        return Tuning(t.dpt, t.chunk, t.n_queues, t.blocks_per_cu, t.cache_dpt, t.kernel, t.cache_kernel), ms.value

    def param_autotune(self, d_cvals: int):
        """Datapoints per lane for the forward and adjoint constant-fitting kernels."""
        f, a = C.c_int32(), C.c_int32()
        self._check(self._L.sr_param_autotune(self._h, d_cvals, C.byref(f), C.byref(a)))
        return f.value, a.value

    def cache_config(self, capacity=2048, min_size=3, min_freq=10):
        self._check(self._L.sr_cache_config(self._h, capacity, min_size, min_freq))

    def cache_policy(self, policy, capacity=0, min_size=4):
        """Select the SR_EVAL_CACHED policy: CACHE_REFERENCE (detect-once, the
        reference's cl_gpu_eval.cu:294-329) or one of the generational
        policies (columns follow the population, stale ones evicted), which
        differ in which subtrees get a column: CACHE_GENERATIONAL every new
        one, CACHE_GENERATIONAL_REUSED one met twice (in one population or in
        two consecutive ones), CACHE_GENERATIONAL_SHARED one met twice in the
        population plus one nested in those that the previous population
        held; or CACHE_DEVICE, planned on the device for the GA population
        (ga_run / ga_evaluate; a subtree executed twice in the population is
        evaluated once and kept).  capacity 0 = sized from free device memory."""
        self._check(self._L.sr_cache_policy(self._h, int(policy), int(capacity), int(min_size)))

    def cache_clear(self):
        self._check(self._L.sr_cache_clear(self._h))

    def cache_keys(self):
        n = C.c_int32()
        self._check(self._L.sr_cache_state(self._h, None, 0, C.byref(n)))
        keys = np.zeros(max(n.value, 1), np.uint64)
        self._check(self._L.sr_cache_state(self._h, keys.ctypes.data, n.value, C.byref(n)))
        return keys[:n.value]

    def cache_last_hints(self):
        n = C.c_int64()
        self._check(self._L.sr_cache_last_hints(self._h, None, 0, C.byref(n)))
        h = np.zeros(max(n.value, 1), np.int32)
        self._check(self._L.sr_cache_last_hints(self._h, h.ctypes.data, n.value, C.byref(n)))
        return h[:n.value]

    def detect(self, pop: Population, min_size=3, min_freq=2, max_subs=1 << 16):
        hints = np.zeros(max(int(pop.offsets[-1] + pop.lengths[-1]) if pop.num_exprs else 1, 1),
                         np.int32)
        flat = np.zeros(max_subs, np.uint64)
        size = np.zeros(max_subs, np.int32)
        count = np.zeros(max_subs, np.int32)
        n = C.c_int32()
        self._check(self._L.sr_detect(self._h, C.byref(pop._c()), min_size, min_freq, hints, flat,
                                      size, count, max_subs, C.byref(n)))
        k = min(n.value, max_subs)
        return {"num_subs": n.value, "flat": flat[:k], "size": size[:k], "count": count[:k],
                "hints": hints}

    # ---- genetic algorithm ------------------------------------------------
    def ga_init(self, params: GaParams, pop: int, seed: int):
        self._check(self._L.sr_ga_init(self._h, C.byref(params), pop, seed))

    def _ga_shape(self):
        pop, cap = C.c_int32(), C.c_int32()
        self._check(self._L.sr_ga_get_population(self._h, None, None, None, C.byref(pop),
                                                 C.byref(cap)))
        return pop.value, cap.value

    def ga_population(self):
        """(tok int32[pop*cap], val float32[pop*cap], len int32[pop], cap)."""
        pop, cap = self._ga_shape()
        tok = np.empty(pop * cap, np.int32)
        val = np.empty(pop * cap, np.float32)
        ln = np.empty(pop, np.int32)
        p, c = C.c_int32(), C.c_int32()
        self._check(self._L.sr_ga_get_population(self._h, tok.ctypes.data, val.ctypes.data,
                                                 ln.ctypes.data, C.byref(p), C.byref(c)))
        return tok, val, ln, cap

    def ga_set_population(self, tok, val, ln, cap):
        self._check(self._L.sr_ga_set_population(
            self._h, np.ascontiguousarray(tok, np.int32), np.ascontiguousarray(val, np.float32),
            np.ascontiguousarray(ln, np.int32), len(ln), cap))

    def ga_evaluate(self, mode=EVAL_BATCHED):
        pop, _ = self._ga_shape()
        m = np.empty(pop, np.float32)
        st = _Stats()
        self._check(self._L.sr_ga_evaluate(self._h, mode, m, C.byref(st)))
        return m, st.as_dict()

    def ga_next(self, sel: SelectParams, fitness, seed: int, gen: int):
        f = None if fitness is None else np.ascontiguousarray(fitness, np.float32)
        self._check(self._L.sr_ga_next(self._h, C.byref(sel), None if f is None else f.ctypes.data,
                                       seed, gen))

    def evolve_step(self, pop: Population, fitness, params: GaParams, sel: SelectParams, seed: int,
                    gen: int):
        """One select + vary step on a host population (sr_evolve_step):
        returns the next generation (tok, val, len, cap), stride cap."""
        E = pop.num_exprs
        cap = params.max_tokens + params.max_depth + 2
        tok = np.zeros(E * cap, np.int32)
        val = np.zeros(E * cap, np.float32)
        ln = np.zeros(E, np.int32)
        f = np.ascontiguousarray(fitness, np.float32)
        assert len(f) == E
        self._check(self._L.sr_evolve_step(self._h, C.byref(pop._c()), f.ctypes.data, C.byref(params),
                                           C.byref(sel), seed, gen, cap, tok.ctypes.data, val.ctypes.data,
                                           ln.ctypes.data))
        return tok, val, ln, cap

    def ga_run(self, sel: SelectParams, mode: int, gens: int, gen0: int, seed: int,
               use_graph: bool = True):
        """gens generations; returns (fitness history [gens, pop], ms, stats)."""
        pop, _ = self._ga_shape()
        hist = np.empty(max(gens, 1) * pop, np.float32)
        ms = C.c_double()
        st = _Stats()
        self._check(self._L.sr_ga_run(self._h, C.byref(sel), mode, gens, gen0, seed,
                                      int(use_graph), hist.ctypes.data, C.byref(ms), C.byref(st)))
        return hist[:gens * pop].reshape(gens, pop), ms.value, st.as_dict()

    # ---- evogp population layout (device tensors) ---------------------------
    def evogp_to_prefix(self, value, type_, size, cap: int):
        """evogp arrays (CUDA tensors [pop, max_len]: float32 value, int16 type,
        int16 subtree size) -> (tok int32 [pop, cap], val float32, len int32)."""
        import torch
        pop, max_len = value.shape
        tok = torch.zeros((pop, cap), dtype=torch.int32, device=value.device)
        val = torch.zeros((pop, cap), dtype=torch.float32, device=value.device)
        ln = torch.zeros(pop, dtype=torch.int32, device=value.device)
        torch.cuda.synchronize(value.device)
        self._check(self._L.sr_evogp_to_prefix(self._h, pop, max_len, value.data_ptr(), type_.data_ptr(),
                                               size.data_ptr(), cap, tok.data_ptr(), val.data_ptr(),
                                               ln.data_ptr()))
        return tok, val, ln

    def prefix_to_evogp(self, tok, val, ln, max_len: int):
        """Stride population (CUDA tensors [pop, cap]) -> evogp arrays."""
        import torch
        pop, cap = tok.shape
        value = torch.zeros((pop, max_len), dtype=torch.float32, device=tok.device)
        type_ = torch.zeros((pop, max_len), dtype=torch.int16, device=tok.device)
        size = torch.zeros((pop, max_len), dtype=torch.int16, device=tok.device)
        torch.cuda.synchronize(tok.device)
        self._check(self._L.sr_prefix_to_evogp(self._h, pop, cap, tok.data_ptr(), val.data_ptr(), ln.data_ptr(),
                                               max_len, value.data_ptr(), type_.data_ptr(), size.data_ptr()))
        return value, type_, size

    def ga_import_evogp(self, value, type_, size):
        import torch
        torch.cuda.synchronize(value.device)
        pop, max_len = value.shape
        self._check(self._L.sr_ga_import_evogp(self._h, pop, max_len, value.data_ptr(), type_.data_ptr(),
                                               size.data_ptr()))

    def ga_export_evogp(self, max_len: int):
        import torch
        pop, _ = self._ga_shape()
        value = torch.zeros((pop, max_len), dtype=torch.float32, device="cuda")
        type_ = torch.zeros((pop, max_len), dtype=torch.int16, device="cuda")
        size = torch.zeros((pop, max_len), dtype=torch.int16, device="cuda")
        torch.cuda.synchronize()
        self._check(self._L.sr_ga_export_evogp(self._h, max_len, value.data_ptr(), type_.data_ptr(),
                                               size.data_ptr()))
        return value, type_, size

    # ---- constant fitting (device pointers) ---------------------------------
    def param_prepare(self, pop: Population, const_slot):
        """Free constants: const_slot[t] = slot 0..7 of token t (CONST tokens) or
        -1.  Returns the initial constant table float32 [E, 8]."""
        init = np.zeros(max(pop.num_exprs, 1) * 8, np.float32)
        self._check(self._L.sr_param_prepare(self._h, C.byref(pop._c()),
                                             np.ascontiguousarray(const_slot, np.int32),
                                             init.ctypes.data))
        return init[:pop.num_exprs * 8].reshape(pop.num_exprs, 8)

    def param_nslots(self):
        n = C.c_int32()
        t = C.c_int64()
        self._check(self._L.sr_param_info(self._h, C.byref(n), None, C.byref(t)))
        out = np.zeros(max(n.value, 1), np.int32)
        self._check(self._L.sr_param_info(self._h, C.byref(n), out.ctypes.data, C.byref(t)))
        return out[:n.value], t.value

    def grad_consts(self, pop: Population, const_mask):
        """dMSE/dconstant for every masked CONST token at the population's own
        values (sr_grad_consts): returns (grad packed like the tokens, mse)."""
        m = np.ascontiguousarray(const_mask, np.uint8)
        g = np.zeros(max(len(m), 1), np.float32)
        mse = np.zeros(max(pop.num_exprs, 1), np.float32)
        self._check(self._L.sr_grad_consts(self._h, C.byref(pop._c()), m.ctypes.data, g.ctypes.data,
                                           mse.ctypes.data))
        return g[:len(m)], mse[:pop.num_exprs]

    def param_eval(self, d_cvals: int, d_mse: int):
        self._check(self._L.sr_param_eval(self._h, d_cvals, d_mse))

    def param_grad(self, d_cvals: int, d_mse: int, d_grad: int):
        self._check(self._L.sr_param_grad(self._h, d_cvals, d_mse, d_grad))

    def param_fd_grad(self, d_cvals: int, h: float, d_grad: int):
        self._check(self._L.sr_param_fd_grad(self._h, d_cvals, h, d_grad))

    def sync(self):
        self._check(self._L.sr_sync(self._h))

    def tree_splice(self, cap, left, lsel, lidx, right, rsel, ridx):
        """Batch subtree replacement; left/right = (tok, val, len) with stride cap."""
        lt, lv, ll = (np.ascontiguousarray(a, d)
                      for a, d in zip(left, (np.int32, np.float32, np.int32)))
        rt, rv, rl = (np.ascontiguousarray(a, d)
                      for a, d in zip(right, (np.int32, np.float32, np.int32)))
        n = len(lsel)
        ot = np.zeros(max(n, 1) * cap, np.int32)
        ov = np.zeros(max(n, 1) * cap, np.float32)
        ol = np.zeros(max(n, 1), np.int32)

        def i32(a):
            return np.ascontiguousarray(a, np.int32)

        self._check(self._L.sr_tree_splice(self._h, n, cap, lt, lv, ll, len(ll), i32(lsel),
                                           i32(lidx), rt, rv, rl, len(rl), i32(rsel), i32(ridx),
                                           ot, ov, ol))
        return ot[:n * cap], ov[:n * cap], ol[:n]
