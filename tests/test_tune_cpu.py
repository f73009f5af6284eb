"""Persistence of the tuning table (sr_tune_save / sr_tune_load): a file's
entries are taken all or none and written back line for line; the file names
the architecture, CU count, variable class and datapoint count of each
entry.  No GPU needed."""
import pytest

from symbolic_regression_gpu_amd import SrError, api

HEADER = "# arch cus xc N dpt chunk n_queues blocks_per_cu cache_dpt kernel cache_kernel\n"


def _entries(path):
    with open(path) as f:
        return sorted(line.strip() for line in f if line.strip() and not line.startswith("#"))


def test_load_then_save_round_trip(tmp_path):
    src = tmp_path / "in.txt"
    lines = ["gfx950 256 4 1048576 3 3072 8 0 3 1 1", "gfx950 256 16 65536 2 6144 8 1 3 0 1",
             "gfx942 304 32 1000 1 3072 1 0 0 1 0"]
    src.write_text(HEADER + "\n".join(lines) + "\n")
    assert api.tune_load(str(src)) == 3
    out = tmp_path / "out.txt"
    api.tune_save(str(out))
    saved = _entries(out)
    for line in lines:
        assert line in saved


@pytest.mark.parametrize("bad", ["gfx950 256 5 1000 3 3072 8 0 3 1 1",     # no such variable class
                                 "gfx950 256 4 1000 9 3072 8 0 3 1 1",     # dpt out of range
                                 "gfx950 256 4 1000 3 3072 4 0 3 1 1",     # queues are 1 or 8
                                 "gfx950 256 4 1000 3 3072 8 0 3 2 1",     # no such kernel
                                 "gfx950 256 4 1000 3 3072"])               # short line
def test_a_bad_line_rejects_the_whole_file(tmp_path, bad):
    src = tmp_path / "bad.txt"
    good = "gfx950 128 4 4242 2 3072 8 0 3 1 1"
    src.write_text(HEADER + good + "\n" + bad + "\n")
    with pytest.raises(SrError):
        api.tune_load(str(src))
    out = tmp_path / "out.txt"
    api.tune_save(str(out))
    assert good not in _entries(out)


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(SrError):
        api.tune_load(str(tmp_path / "absent.txt"))
