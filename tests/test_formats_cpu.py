"""On-disk formats (digest gen_k.txt and row-major data files) against the
reference's own parsers.

tests/golden/formats.npz holds parse_input_info (utils.cpp:122-233) and
load_data_file (utils.cpp:235-300) outputs, produced by the compiled
reference (tests/golden/make_golden.py), for the reference's example files
(data/examples/, copied under tests/golden/formats/) and two synthetic files
with awkward number spellings (signed zero, 1.5e-300, 19-digit pi, E+ exponents,
tabs, CRLF rows, a short read).  Our readers must match bit for bit; writers
must round-trip exactly.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from symbolic_regression_gpu_amd.api import (Population, SrError, read_data_file, read_digest,
                                             write_data_file, write_digest)
from tests.helpers import GRAMMAR_BINARY, GRAMMAR_UNARY, random_population

HERE = os.path.dirname(os.path.abspath(__file__))
FDIR = os.path.join(HERE, "golden", "formats")
G = np.load(os.path.join(HERE, "golden", "formats.npz"))


@pytest.mark.parametrize("tag,name", [("sample", "sample_input.txt"), ("synth", "synth_digest.txt")])
def test_digest_matches_reference_parser(tag, name):
    d = read_digest(os.path.join(FDIR, name))
    assert np.array_equal(d.num_vars, G[f"{tag}_nv"])
    assert np.array_equal(d.num_dps, G[f"{tag}_ndps"])
    assert np.array_equal(d.population.lengths, G[f"{tag}_ntok"])
    assert np.array_equal(d.population.tokens, G[f"{tag}_tok"])
    assert np.array_equal(d.population.values.view(np.uint64), G[f"{tag}_val"].view(np.uint64))
    assert d.data_paths == list(G[f"{tag}_paths"])
    assert d.shared == bool(G[f"{tag}_shared"])
    # the packed layout is the evaluator's population layout
    off = np.concatenate([[0], np.cumsum(d.population.lengths)[:-1]])
    assert np.array_equal(d.population.offsets, off)


def test_synthetic_digest_special_values():
    d = read_digest(os.path.join(FDIR, "synth_digest.txt"))
    v = d.population.values
    assert np.signbit(v[5]) and v[5] == 0.0  # "-0.0"
    assert v[8] == 1.5e-300
    assert v[10] == np.float64("3.141592653589793238") and v[11] == -2.5e7


@pytest.mark.parametrize("tag,name,nv,ndps", [("s1", "sample_1", 1, 1000), ("s2", "sample_2", 3, 1000),
                                              ("sd", "synth_data.txt", 2, 64),
                                              ("short", "synth_data.txt", 2, 70)])
def test_data_file_matches_reference_parser(tag, name, nv, ndps):
    cols, parsed = read_data_file(os.path.join(FDIR, name), nv, ndps)
    want = G[f"data_{tag}"]
    assert cols.shape == want.shape
    assert np.array_equal(cols.view(np.uint64), want.view(np.uint64))
    rows = 64 if tag in ("sd", "short") else ndps
    assert parsed == min(rows, ndps) * (nv + 1)


def test_shared_flag_rules(tmp_path):
    pop = Population.from_exprs([([1, -1, 0], [0, 0, 2.0]), ([10, -1], [0, 1])])
    one = Population.from_exprs([([10, -1], [0, 1])])
    write_digest(str(tmp_path / "a.txt"), pop, 2, 100, "data.txt")
    assert read_digest(str(tmp_path / "a.txt")).shared
    write_digest(str(tmp_path / "b.txt"), pop, 2, 100, ["data.txt", "other.txt"])
    assert not read_digest(str(tmp_path / "b.txt")).shared
    write_digest(str(tmp_path / "c.txt"), pop, [2, 3], 100, "data.txt")
    assert not read_digest(str(tmp_path / "c.txt")).shared
    # one expression is never "shared" (utils.cpp:202-223)
    write_digest(str(tmp_path / "d.txt"), one, 2, 100, "data.txt")
    assert not read_digest(str(tmp_path / "d.txt")).shared


def test_round_trips(tmp_path):
    pop = random_population(50, 5, nv=3, max_tokens=40, max_depth=6, unary=GRAMMAR_UNARY,
                            binary=GRAMMAR_BINARY)
    api = Population(pop.tok, pop.val, pop.off, pop.len)
    p = str(tmp_path / "gen_0.txt")
    write_digest(p, api, 3, 1234, "shared_data.txt")
    d = read_digest(p)
    assert np.array_equal(d.population.tokens, pop.tok)
    assert np.array_equal(d.population.values.view(np.uint64), pop.val.view(np.uint64))
    assert np.array_equal(d.population.lengths, pop.len)
    assert d.shared and d.data_paths == ["shared_data.txt"] * 50
    # the reference's own parser reads what we write
    if os.path.exists(O.REF_PATH):
        R = O.ref()
        import ctypes as C
        nv = np.zeros(64, np.int32)
        tok = np.zeros(4096, np.int32)
        val = np.zeros(4096, np.float64)
        paths = C.create_string_buffer(64 * 256)
        sh = C.c_int()
        E = R.ref_parse_digest(p.encode(), 64, 4096, 256, nv.ctypes.data, np.zeros(64, np.int32).ctypes.data,
                               np.zeros(64, np.int32).ctypes.data, tok.ctypes.data, val.ctypes.data, paths,
                               C.byref(sh))
        assert E == 50 and sh.value == 1
        T = int(pop.len.sum())
        assert np.array_equal(tok[:T], pop.tok) and np.array_equal(val[:T], pop.val)
    rng = np.random.default_rng(1)
    cols = rng.normal(size=(4, 300)) * np.array([[1], [1e-30], [1e30], [1]])
    cols[0, 0] = -0.0
    q = str(tmp_path / "shared_data.txt")
    write_data_file(q, cols)
    back, n = read_data_file(q, 3, 300)
    assert n == 1200 and np.array_equal(back.view(np.uint64), cols.view(np.uint64))


def test_errors(tmp_path):
    with pytest.raises(SrError) as e:
        read_digest(str(tmp_path / "missing.txt"))
    assert e.value.code == -9
    with pytest.raises(SrError):
        read_data_file(str(tmp_path / "missing.txt"), 1, 10)
    for text in ("0\n", "-3\n", "x\n", "2\n1 10 3\n1 -1 0\n0 0 1.0\ndata\n1 10 2\n10 -1\n"):
        f = tmp_path / "bad.txt"
        f.write_text(text)
        with pytest.raises(SrError) as e:
            read_digest(str(f))
        assert e.value.code == -3, text


def test_generation_row_restates_sr_system_statistics():
    """cl_sr_system.cpp:229-283: best over finite MSEs (1e9 when none),
    median = sorted finite [n // 2] (0 when none), covered = total x coverage."""
    from symbolic_regression_gpu_amd.replay import CSV_HEADER, generation_row
    st = {"coverage": 0.25, "total_eval_time_ms": 3.0, "detect_time_ms": 1.0, "transfer_time_ms": 0.5,
          "jit_compile_time_ms": 0.0, "kernel_time_ms": 1.5, "num_subtrees": 7, "avg_subtree_size": 4.5}
    r = generation_row(3, [5.0, np.nan, 1.0, np.inf, 3.0, 2.0], [4, 5, 6, 7, 8, 10], st)
    assert list(r) == CSV_HEADER.split(",")
    assert r["BestMSE"] == 1.0 and r["MedianMSE"] == 3.0  # sorted finite [1, 2, 3, 5], index 2
    assert r["TotalTokens"] == 40 and r["CoveredTokens"] == 10.0
    assert (r["TotalTimeMs"], r["DetectTimeMs"], r["H2D_D2H_TimeMs"], r["KernelTimeMs"]) == (3.0, 1.0, 0.5, 1.5)
    none = generation_row(0, [np.nan, np.inf], [3, 3], {})
    assert none["BestMSE"] == 1e9 and none["MedianMSE"] == 0.0 and none["Coverage"] == 0.0
