"""The C-ABI library loads, exports every function include/sr_hip.h declares,
and reports errors as status codes (no GPU needed)."""
import ctypes as C
import os
import re

import pytest

from symbolic_regression_gpu_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "sr_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sr_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(api.LIB_PATH)
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_null_arguments_are_status_codes():
    L = api.lib()
    assert L.sr_ctx_create(0, None) == -1
    assert L.sr_ctx_destroy(None) == 0
    raw = C.CDLL(api.LIB_PATH)  # untyped symbols: null pointers pass through ctypes
    assert raw.sr_set_data(None, None, 0, C.c_int64(0)) == -1
    assert raw.sr_eval_mse(None, None, 0, None, None, None) == -1
    assert raw.sr_ga_run(None, None, 0, 0, 0, C.c_uint64(0), 0, None, None, None) == -1


def test_no_device_is_a_status_code():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU device node is present")
    h = C.c_void_p()
    assert api.lib().sr_ctx_create(0, C.byref(h)) == -8


def test_compile_hook_runs_without_gpu():
    code, arg, need = api.compile_expr([1, -1, 0], [0, 0, 2.0], nv=1)
    assert len(code) == 2 and need == 0


def test_one_hip_runtime_per_process():
    """A process that loads the library through the package runs one HIP
    runtime (torch's libamdhip64.so.7, which the package loads first, serves
    the library too), of the same major version as the headers the library
    was built with.  Checked in a fresh interpreter: loading libsrhip.so
    directly before torch maps the one its rpath names as well."""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from symbolic_regression_gpu_amd import api\n"
            "api.lib()\n"
            "paths = sorted({l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l})\n"
            "print(len(paths), *api.runtime_info())\n" % ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-500:]
    n, built, runtime, driver, same = r.stdout.split()[-5:]
    assert n == "1" and same == "True" and int(built) // 10_000_000 == 7, r.stdout
