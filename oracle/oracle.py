"""ctypes front end of the CPU oracle (oracle/sr_oracle.c) and, when it has been
built in this container, of the reference harness (oracle/_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the symbolic_regression_gpu_amd package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class GaParams(C.Structure):
    _fields_ = [
        ("num_vars", C.c_int), ("max_tokens", C.c_int), ("max_depth", C.c_int),
        ("const_min", C.c_float), ("const_max", C.c_float),
        ("prob_leaf", C.c_float), ("prob_const_leaf", C.c_float), ("prob_unary", C.c_float),
        ("n_unary", C.c_int), ("n_binary", C.c_int),
        ("unary_ops", C.c_int * 32), ("binary_ops", C.c_int * 32),
    ]


class SelectParams(C.Structure):
    _fields_ = [("tournament", C.c_int), ("p_mutation", C.c_float), ("elite_frac", C.c_double)]


class DetectResult(C.Structure):
    _fields_ = [
        ("num_subs", C.c_int),
        ("sub_expr", C.POINTER(C.c_int)), ("sub_start", C.POINTER(C.c_int)),
        ("sub_size", C.POINTER(C.c_int)), ("sub_hash", C.POINTER(C.c_uint64)),
        ("occ_offset", C.POINTER(C.c_int)), ("occ_expr", C.POINTER(C.c_int)),
        ("occ_start", C.POINTER(C.c_int)), ("hints", C.POINTER(C.c_int)),
        ("covered_tokens", C.c_int64), ("avg_subtree_size", C.c_double),
        ("coverage", C.c_double),
    ]


def build(ref: bool = False) -> None:
    subprocess.run(["make", "-s", "-C", HERE] + (["ref"] if ref else []), check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.sro_op32.restype = C.c_float
        L.sro_op32.argtypes = [C.c_int, C.c_float, C.c_float]
        L.sro_op64.restype = C.c_double
        L.sro_op64.argtypes = [C.c_int, C.c_double, C.c_double]
        L.sro_predict32.argtypes = [_i32p, _f64p, C.c_int, _f32p, C.c_int, C.c_int64, _f32p]
        L.sro_predict64.argtypes = [_i32p, _f64p, C.c_int, _f32p, C.c_int, C.c_int64, _f64p]
        L.sro_predict64_bound.argtypes = [_i32p, _f64p, C.c_int, _f32p, C.c_int, C.c_int64, _f64p,
                                          _f64p]
        L.sro_mse_grad64_bound.argtypes = [_i32p, _f64p, _i32p, C.c_int, _f32p, C.c_int, C.c_int64,
                                           _f32p, C.c_int, C.POINTER(C.c_double), _f64p, _f64p,
                                           C.POINTER(C.c_int64)]
        L.sro_mse32.argtypes = [_i32p, _f64p, _i32p, _i32p, C.c_int, _f32p, C.c_int, C.c_int64,
                                C.c_int, _f32p, _i64p]
        L.sro_mse64.argtypes = [_i32p, _f64p, _i32p, _i32p, C.c_int, _f32p, C.c_int, C.c_int64,
                                C.c_int, _f64p]
        L.sro_stack_depth.argtypes = [_i32p, C.c_int]
        L.sro_subtree_size.argtypes = [_i32p, C.c_int, C.c_int]
        L.sro_hash_flat.restype = C.c_uint64
        L.sro_hash_flat.argtypes = [_i32p, _f64p, C.c_int]
        L.sro_subtree_merkle.argtypes = [_i32p, _f64p, C.c_int, _u64p, _i32p]
        L.sro_detect.argtypes = [_i32p, _f64p, _i32p, _i32p, C.c_int, C.c_int, C.c_int,
                                 C.POINTER(DetectResult)]
        L.sro_detect_free.argtypes = [C.POINTER(DetectResult)]
        L.sro_cache_keys_create.restype = C.c_void_p
        L.sro_cache_keys_create.argtypes = [C.c_int]
        L.sro_cache_keys_free.argtypes = [C.c_void_p]
        L.sro_cache_keys_get.restype = C.c_int
        L.sro_cache_keys_get.argtypes = [C.c_void_p, C.c_void_p]
        L.sro_cache_update.restype = C.c_int
        L.sro_cache_update.argtypes = [C.POINTER(DetectResult), _i32p, _f64p, _i32p, C.c_void_p, _i32p, _i32p]
        L.sro_mark_cached.argtypes = [_i32p, _f64p, _i32p, _i32p, C.c_int, C.c_void_p, C.c_int, _i32p]
        L.sro_skeleton.restype = C.c_int
        L.sro_skeleton.argtypes = [_i32p, _f64p, _i32p, _i32p, C.c_int, _i32p, _i32p, _f64p,
                                   _i32p, _i32p]
        L.sro_predict32_cached.argtypes = [_i32p, _f64p, C.c_int, _f32p, C.c_int, C.c_int64,
                                           _f32p, _f32p]
        L.sro_philox4x32_10.argtypes = [np.ctypeslib.ndpointer(dtype=np.uint32)] * 3
        L.sro_ga_params_default.argtypes = [C.POINTER(GaParams)]
        L.sro_ga_next_generation.argtypes = [
            C.POINTER(GaParams), C.POINTER(SelectParams), C.c_int, C.c_int, _i32p, _f32p, _i32p,
            _f32p, C.c_uint64, C.c_uint32, _i32p, _f32p, _i32p]
        L.sro_ga_init_population.argtypes = [C.POINTER(GaParams), C.c_int, C.c_int, C.c_uint64,
                                             _i32p, _f32p, _i32p]
        L.sro_fitness_order.argtypes = [_f32p, C.c_int, _i32p]
        _vp = C.c_void_p
        L.sro_evogp_to_prefix.argtypes = [C.c_int, C.c_int, _vp, _vp, _vp, C.c_int, _vp, _vp, _vp]
        L.sro_prefix_to_evogp.argtypes = [C.c_int, C.c_int, _vp, _vp, _vp, C.c_int, _vp, _vp, _vp]
        L.sro_ga_subtree_size.argtypes = [_i32p, C.c_int, C.c_int]
        L.sro_subtree_predict64.restype = C.c_int
        L.sro_subtree_predict64.argtypes = [_i32p, _f64p, _i32p, _i32p, C.c_int, _f64p, C.c_int, C.c_int64,
                                            C.c_int, _f64p, _vp]
        L.sro_subtree_state_create.restype = _vp
        L.sro_subtree_state_free.argtypes = [_vp]
        L.sro_subtree_state_size.restype = C.c_int
        L.sro_subtree_state_size.argtypes = [_vp]
        L.sro_subtree_state_predict64.restype = C.c_int
        L.sro_subtree_state_predict64.argtypes = [_vp, _i32p, _f64p, _i32p, _i32p, C.c_int, _f64p, C.c_int,
                                                  C.c_int64, C.c_int, _f64p, _vp]
        _bind_ga_hooks(L)
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """Reference harness (only in a container holding the reference tree)."""
    global _ref
    if _ref is None:
        R = C.CDLL(REF_PATH)
        R.ref_op32.restype = C.c_float
        R.ref_op32.argtypes = [C.c_int, C.c_float, C.c_float]
        R.ref_op64.restype = C.c_double
        R.ref_op64.argtypes = [C.c_int, C.c_double, C.c_double]
        R.ref_eval_point32.restype = C.c_float
        R.ref_eval_point32.argtypes = [_i32p, _f64p, C.c_int, _f32p, C.c_int]
        R.ref_mse32.argtypes = [_i32p, _f64p, _i32p, _i32p, C.c_int, _f32p, C.c_int, C.c_int64,
                                _f32p]
        R.ref_parse_digest.restype = C.c_int
        R.ref_parse_digest.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 5 + \
            [C.c_char_p, C.POINTER(C.c_int)]
        R.ref_load_data.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_void_p]
        R.ref_hash_flat.restype = C.c_uint64
        R.ref_hash_flat.argtypes = [_i32p, _f64p, C.c_int]
        R.ref_detect.restype = C.c_int
        R.ref_detect.argtypes = [_i32p, _f64p, _i32p, _i32p, C.c_int, C.c_int, C.c_int, _i32p,
                                 _i32p, _u64p, _i32p, C.c_int]
        R.ref_ga_op_log.restype = C.c_int
        R.ref_ga_op_log.argtypes = [C.c_int, C.c_int, C.POINTER(GaParams), _i32p, _f32p, C.c_int,
                                    _i32p, _f32p, C.c_int, _i32p, _f32p, _i32p, _f32p, _f32p, _i32p,
                                    C.c_int, C.POINTER(C.c_int)]
        R.ref_subtree_predict64.argtypes = [_i32p, _f64p, _i32p, _i32p, C.c_int, _f64p, C.c_int, C.c_int,
                                            _f64p]
        R.ref_subtree_cache_create.restype = C.c_void_p
        R.ref_subtree_cache_free.argtypes = [C.c_void_p]
        R.ref_subtree_stateful_predict64.restype = C.c_int
        R.ref_subtree_stateful_predict64.argtypes = [C.c_void_p, _i32p, _f64p, _i32p, _i32p, C.c_int, _f64p,
                                                     C.c_int, C.c_int, _f64p]
        R.ref_ga_op.restype = C.c_int
        R.ref_ga_op.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(GaParams), _i32p, _f32p,
                                C.c_int, _i32p, _f32p, C.c_int, _i32p, _f32p]
        _ref = R
    return _ref


# --------------------------------------------------------------------------
# numpy-level helpers
# --------------------------------------------------------------------------

class Population:
    """Packed prefix-token population: expression e is tok[off[e]:off[e]+len[e]]."""

    def __init__(self, exprs):
        toks, vals, offs, lens = [], [], [], []
        n = 0
        for t, v in exprs:
            offs.append(n)
            lens.append(len(t))
            toks.extend(int(x) for x in t)
            vals.extend(float(x) for x in v)
            n += len(t)
        self.tok = np.asarray(toks, dtype=np.int32)
        self.val = np.asarray(vals, dtype=np.float64)
        self.off = np.asarray(offs, dtype=np.int32)
        self.len = np.asarray(lens, dtype=np.int32)

    @property
    def E(self):
        return int(self.len.shape[0])

    def expr(self, e):
        a, b = int(self.off[e]), int(self.off[e] + self.len[e])
        return self.tok[a:b], self.val[a:b]


def op32(op, a, b=0.0):
    return lib().sro_op32(op, a, b)


def op64(op, a, b=0.0):
    return lib().sro_op64(op, a, b)


def predict32(tok, val, X, nv, N):
    out = np.empty(N, dtype=np.float32)
    lib().sro_predict32(np.ascontiguousarray(tok, np.int32), np.ascontiguousarray(val, np.float64),
                        len(tok), np.ascontiguousarray(X, np.float32), nv, N, out)
    return out


def predict64(tok, val, X, nv, N):
    out = np.empty(N, dtype=np.float64)
    lib().sro_predict64(np.ascontiguousarray(tok, np.int32), np.ascontiguousarray(val, np.float64),
                        len(tok), np.ascontiguousarray(X, np.float32), nv, N, out)
    return out


def predict64_bound(tok, val, X, nv, N):
    """FP64 values and first-order error bounds of an fp32 evaluation."""
    out = np.empty(N, dtype=np.float64)
    bnd = np.empty(N, dtype=np.float64)
    lib().sro_predict64_bound(np.ascontiguousarray(tok, np.int32),
                              np.ascontiguousarray(val, np.float64), len(tok),
                              np.ascontiguousarray(X, np.float32), nv, N, out, bnd)
    return out, bnd


def mse_grad64_bound(tok, val, slot, X, nv, N, cvals, lane_terms):
    """FP64 MSE, dMSE/dconstants [8] and the error bound [8] of an fp32
    adjoint evaluation (inf where a branch can flip), and the valid count."""
    m = C.c_double()
    cnt = C.c_int64()
    g = np.empty(8, dtype=np.float64)
    b = np.empty(8, dtype=np.float64)
    lib().sro_mse_grad64_bound(np.ascontiguousarray(tok, np.int32), np.ascontiguousarray(val, np.float64),
                               np.ascontiguousarray(slot, np.int32), len(tok),
                               np.ascontiguousarray(X, np.float32), nv, N,
                               np.ascontiguousarray(cvals, np.float32), lane_terms, C.byref(m), g, b,
                               C.byref(cnt))
    return m.value, g, b, cnt.value


def mse32(pop: Population, X, nv, N, threads=1):
    out = np.empty(pop.E, dtype=np.float32)
    valid = np.empty(pop.E, dtype=np.int64)
    lib().sro_mse32(pop.tok, pop.val, pop.off, pop.len, pop.E, np.ascontiguousarray(X, np.float32),
                    nv, N, threads, out, valid)
    return out, valid


def mse64(pop: Population, X, nv, N, threads=1):
    out = np.empty(pop.E, dtype=np.float64)
    lib().sro_mse64(pop.tok, pop.val, pop.off, pop.len, pop.E, np.ascontiguousarray(X, np.float32),
                    nv, N, threads, out)
    return out


def stack_depth(tok):
    t = np.ascontiguousarray(tok, np.int32)
    return lib().sro_stack_depth(t, len(t))


def hash_flat(tok, val):
    t = np.ascontiguousarray(tok, np.int32)
    return int(lib().sro_hash_flat(t, np.ascontiguousarray(val, np.float64), len(t)))


def merkle(tok, val):
    t = np.ascontiguousarray(tok, np.int32)
    h = np.empty(len(t), dtype=np.uint64)
    s = np.empty(len(t), dtype=np.int32)
    lib().sro_subtree_merkle(t, np.ascontiguousarray(val, np.float64), len(t), h, s)
    return h, s


def detect(pop: Population, min_size=3, min_freq=2):
    r = DetectResult()
    lib().sro_detect(pop.tok, pop.val, pop.off, pop.len, pop.E, min_size, min_freq, C.byref(r))
    n = r.num_subs
    T = int(pop.len.sum())
    occ_off = [r.occ_offset[i] for i in range(n + 1)]
    out = {
        "num_subs": n,
        "sub_expr": np.array([r.sub_expr[i] for i in range(n)], np.int32),
        "sub_start": np.array([r.sub_start[i] for i in range(n)], np.int32),
        "sub_size": np.array([r.sub_size[i] for i in range(n)], np.int32),
        "sub_hash": np.array([r.sub_hash[i] for i in range(n)], np.uint64),
        "occurrences": [
            [(r.occ_expr[k], r.occ_start[k]) for k in range(occ_off[i], occ_off[i + 1])]
            for i in range(n)
        ],
        "hints": np.array([r.hints[i] for i in range(T)], np.int32),
        "covered_tokens": r.covered_tokens,
        "avg_subtree_size": r.avg_subtree_size,
        "coverage": r.coverage,
    }
    lib().sro_detect_free(C.byref(r))
    out["sub_flat"] = np.array(
        [hash_flat(*_span(pop, out["sub_expr"][i], out["sub_start"][i], out["sub_size"][i]))
         for i in range(n)], np.uint64)
    return out


def _span(pop, e, start, size):
    a = int(pop.off[e]) + int(start)
    return pop.tok[a:a + size], pop.val[a:a + size]


class CacheKeys:
    """Slots of the cross-generation subtree cache (cl_gpu_eval.cu:384-397):
    flat hash + body per slot, bodies compared on lookup."""

    def __init__(self, capacity):
        self._k = lib().sro_cache_keys_create(int(capacity))

    def update(self, pop: Population, min_size=3, min_freq=10):
        """Detect, then give new subtrees slots; returns (new_sub, new_slot)."""
        det = DetectResult()
        L = lib()
        L.sro_detect(pop.tok, pop.val, pop.off, pop.len, pop.E, min_size, min_freq, C.byref(det))
        try:
            new_sub = np.zeros(max(det.num_subs, 1), np.int32)
            new_slot = np.zeros(max(det.num_subs, 1), np.int32)
            n = L.sro_cache_update(C.byref(det), pop.tok, pop.val, pop.off, self._k, new_sub, new_slot)
        finally:
            L.sro_detect_free(C.byref(det))
        return new_sub[:n], new_slot[:n]

    @property
    def keys(self):
        n = lib().sro_cache_keys_get(self._k, None)
        out = np.zeros(max(n, 1), np.uint64)
        lib().sro_cache_keys_get(self._k, out.ctypes.data)
        return out[:n]

    def mark(self, pop: Population, min_size=3):
        hints = np.empty(max(int(pop.len.sum()), 1), dtype=np.int32)
        lib().sro_mark_cached(pop.tok, pop.val, pop.off, pop.len, pop.E, self._k, min_size, hints)
        return hints[:int(pop.len.sum())]

    def close(self):
        if self._k:
            lib().sro_cache_keys_free(self._k)
            self._k = None

    __del__ = close


def skeleton(pop: Population, hints):
    T = int(pop.len.sum())
    sk_tok = np.empty(max(T, 1), np.int32)
    sk_val = np.empty(max(T, 1), np.float64)
    sk_off = np.empty(pop.E, np.int32)
    sk_len = np.empty(pop.E, np.int32)
    n = lib().sro_skeleton(pop.tok, pop.val, pop.off, pop.len, pop.E,
                           np.ascontiguousarray(hints, np.int32), sk_tok, sk_val, sk_off, sk_len)
    return sk_tok[:n], sk_val[:n], sk_off, sk_len


def predict32_cached(tok, val, X, nv, N, cache):
    out = np.empty(N, dtype=np.float32)
    lib().sro_predict32_cached(np.ascontiguousarray(tok, np.int32),
                               np.ascontiguousarray(val, np.float64), len(tok),
                               np.ascontiguousarray(X, np.float32), nv, N,
                               np.ascontiguousarray(cache, np.float32), out)
    return out


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    o = np.empty(4, np.uint32)
    lib().sro_philox4x32_10(c, k, o)
    return o


def ga_params(**kw) -> GaParams:
    p = GaParams()
    lib().sro_ga_params_default(C.byref(p))
    for k, v in kw.items():
        if k in ("unary_ops", "binary_ops"):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = int(x)
            setattr(p, "n_unary" if k == "unary_ops" else "n_binary", len(v))
        else:
            setattr(p, k, v)
    return p


def ga_init_population(p: GaParams, pop, cap, seed):
    tok = np.zeros(pop * cap, np.int32)
    val = np.zeros(pop * cap, np.float32)
    ln = np.zeros(pop, np.int32)
    rc = lib().sro_ga_init_population(C.byref(p), pop, cap, seed, tok, val, ln)
    return rc, tok, val, ln


def ga_next_generation(p: GaParams, tournament, p_mutation, elite_frac, pop, cap, tok, val, ln,
                       fit, seed, gen):
    sp = SelectParams(tournament, p_mutation, elite_frac)
    otok = np.zeros(pop * cap, np.int32)
    oval = np.zeros(pop * cap, np.float32)
    oln = np.zeros(pop, np.int32)
    rc = lib().sro_ga_next_generation(
        C.byref(p), C.byref(sp), pop, cap, np.ascontiguousarray(tok, np.int32),
        np.ascontiguousarray(val, np.float32), np.ascontiguousarray(ln, np.int32),
        np.ascontiguousarray(fit, np.float32), seed, gen, otok, oval, oln)
    return rc, otok, oval, oln


UNIFORM_FN = C.CFUNCTYPE(C.c_float, C.c_void_p, C.c_float, C.c_float)
RANDINT_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int)


class RngHooks(C.Structure):
    _fields_ = [("uniform", UNIFORM_FN), ("randint", RANDINT_FN), ("state", C.c_void_p)]


def ga_op_with_hooks(op, p: GaParams, a, b, uniform_py, randint_py):
    """Run the oracle's generate (op 0) / mutate (1) / crossover (2) with Python
    RNG hooks; a, b = (tok int32, val float32) parents."""
    L = lib()
    hooks = RngHooks(UNIFORM_FN(lambda _s, lo, hi: uniform_py(lo, hi)),
                     RANDINT_FN(lambda _s, lo, hi: randint_py(lo, hi)), None)
    tok = np.zeros(4096, np.int32)
    val = np.zeros(4096, np.float32)
    at = np.ascontiguousarray(a[0], np.int32)
    av = np.ascontiguousarray(a[1], np.float32)
    bt = np.ascontiguousarray(b[0], np.int32)
    bv = np.ascontiguousarray(b[1], np.float32)
    if op == 0:
        n = L.sro_ga_generate(C.byref(p), p.max_tokens, C.byref(hooks), tok, val)
    elif op == 1:
        n = L.sro_ga_mutate(C.byref(p), at, av, len(at), C.byref(hooks), tok, val)
    else:
        n = L.sro_ga_crossover(C.byref(p), at, av, len(at), bt, bv, len(bt), C.byref(hooks), tok, val)
    return tok[:n].copy(), val[:n].copy()


class _PhiloxStreamC(C.Structure):
    _fields_ = [("key", C.c_uint32 * 2), ("c1", C.c_uint32), ("c2", C.c_uint32), ("c3", C.c_uint32),
                ("word", C.c_uint32)]


class PhiloxStream:
    """The oracle's Philox4x32-10 stream (seed, stream, gen, domain) behind its
    own sro_rng binding: uniform(lo, hi) / randint(lo, hi) as the GA draws."""

    def __init__(self, seed, stream, gen, domain):
        L = lib()
        self._st = _PhiloxStreamC()
        self._r = RngHooks()
        L.sro_philox_stream_init(C.byref(self._st), C.c_uint64(seed), C.c_uint32(stream), C.c_uint32(gen),
                                 C.c_uint32(domain))
        L.sro_rng_bind_philox(C.byref(self._r), C.byref(self._st))

    def uniform(self, lo, hi):
        return self._r.uniform(C.addressof(self._r), lo, hi)

    def randint(self, lo, hi):
        return self._r.randint(C.addressof(self._r), lo, hi)


CHILD_RNG_FN = C.CFUNCTYPE(C.c_void_p, C.c_void_p, C.c_int)


def ga_next_generation_hooked(p: GaParams, tournament, p_mutation, elite_frac, pop, cap, tok, val, ln, fit,
                              rng_for_child):
    """sro_ga_next_generation_hooked: rng_for_child(c) returns (uniform_py,
    randint_py) for child c.  Returns (rc, tok, val, len, parent1, parent2)."""
    L = lib()
    sp = SelectParams(tournament, p_mutation, elite_frac)
    keep = []

    def child(_ud, c):
        u, r = rng_for_child(c)
        h = RngHooks(UNIFORM_FN(lambda _s, lo, hi: u(lo, hi)), RANDINT_FN(lambda _s, lo, hi: r(lo, hi)), None)
        keep.append(h)
        return C.addressof(h)

    cb = CHILD_RNG_FN(child)
    otok = np.zeros(pop * cap, np.int32)
    oval = np.zeros(pop * cap, np.float32)
    oln = np.zeros(pop, np.int32)
    p1 = np.zeros(pop, np.int32)
    p2 = np.zeros(pop, np.int32)
    rc = L.sro_ga_next_generation_hooked(
        C.byref(p), C.byref(sp), pop, cap, np.ascontiguousarray(tok, np.int32),
        np.ascontiguousarray(val, np.float32), np.ascontiguousarray(ln, np.int32),
        np.ascontiguousarray(fit, np.float32), cb, None, otok, oval, oln, p1, p2)
    return rc, otok, oval, oln, p1, p2


def _bind_ga_hooks(L):
    L.sro_philox_stream_init.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]
    L.sro_rng_bind_philox.argtypes = [C.c_void_p, C.c_void_p]
    L.sro_ga_next_generation_hooked.argtypes = [
        C.POINTER(GaParams), C.POINTER(SelectParams), C.c_int, C.c_int, _i32p, _f32p, _i32p, _f32p,
        CHILD_RNG_FN, C.c_void_p, _i32p, _f32p, _i32p, _i32p, _i32p]
    L.sro_ga_generate.argtypes = [C.POINTER(GaParams), C.c_int, C.POINTER(RngHooks), _i32p, _f32p]
    L.sro_ga_mutate.argtypes = [C.POINTER(GaParams), _i32p, _f32p, C.c_int, C.POINTER(RngHooks),
                                _i32p, _f32p]
    L.sro_ga_crossover.argtypes = [C.POINTER(GaParams), _i32p, _f32p, C.c_int, _i32p, _f32p, C.c_int,
                                   C.POINTER(RngHooks), _i32p, _f32p]


def evogp_to_prefix(value, type_, size, max_len, cap):
    """evogp arrays [pop * max_len] -> (tok, val [pop * cap], len [pop])."""
    pop = len(value) // max_len
    tok = np.zeros(pop * cap, np.int32)
    val = np.zeros(pop * cap, np.float32)
    ln = np.zeros(pop, np.int32)
    lib().sro_evogp_to_prefix(pop, max_len, np.ascontiguousarray(value, np.float32).ctypes.data,
                              np.ascontiguousarray(type_, np.int16).ctypes.data,
                              np.ascontiguousarray(size, np.int16).ctypes.data, cap, tok.ctypes.data,
                              val.ctypes.data, ln.ctypes.data)
    return tok, val, ln


def prefix_to_evogp(tok, val, ln, cap, max_len):
    """Stride-cap prefix population -> evogp arrays [pop * max_len]."""
    pop = len(ln)
    value = np.zeros(pop * max_len, np.float32)
    type_ = np.zeros(pop * max_len, np.int16)
    size = np.zeros(pop * max_len, np.int16)
    lib().sro_prefix_to_evogp(pop, cap, np.ascontiguousarray(tok, np.int32).ctypes.data,
                              np.ascontiguousarray(val, np.float32).ctypes.data,
                              np.ascontiguousarray(ln, np.int32).ctypes.data, max_len, value.ctypes.data,
                              type_.ctypes.data, size.ctypes.data)
    return value, type_, size


def _x64(X, nv, N):
    X = np.ascontiguousarray(X, np.float64).reshape(-1)
    assert X.size == (nv + 1) * N
    return X


def subtree_predict64(pop: Population, X, nv, N, threads=8, want_hints=False):
    """CPU subtree evaluator, FP64 (cpu_subtree.cpp:135-340).  X: [(nv+1), N]
    (widened to double).  Returns (pred [E, N], number of detected subtrees
    [, hint map])."""
    pred = np.empty((pop.E, N), np.float64)
    hints = np.empty(max(len(pop.tok), 1), np.int32) if want_hints else None
    S = lib().sro_subtree_predict64(pop.tok, pop.val, pop.off, pop.len, pop.E, _x64(X, nv, N), nv, N, threads,
                                    pred, None if hints is None else hints.ctypes.data)
    return (pred, S, hints) if want_hints else (pred, S)


class SubtreeState:
    """Stateful CPU subtree evaluator, FP64 (cpu_subtree.cpp:445-501): a cache
    of subtree columns kept across calls on one dataset."""

    def __init__(self):
        self._s = lib().sro_subtree_state_create()

    def predict64(self, pop: Population, X, nv, N, threads=8):
        """-> (pred [E, N], new cache entries, hint map packed like pop.tok)."""
        pred = np.empty((pop.E, N), np.float64)
        hints = np.empty(max(len(pop.tok), 1), np.int32)
        new = lib().sro_subtree_state_predict64(self._s, pop.tok, pop.val, pop.off, pop.len, pop.E,
                                                _x64(X, nv, N), nv, N, threads, pred, hints.ctypes.data)
        return pred, new, hints

    @property
    def size(self):
        return lib().sro_subtree_state_size(self._s)

    def close(self):
        if self._s:
            lib().sro_subtree_state_free(self._s)
            self._s = None

    __del__ = close


def fitness_order(fit):
    f = np.ascontiguousarray(fit, np.float32)
    o = np.empty(len(f), np.int32)
    lib().sro_fitness_order(f, len(f), o)
    return o
