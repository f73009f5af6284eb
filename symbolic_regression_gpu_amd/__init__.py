"""MI355X-native symbolic-regression evaluator (HIP kernels behind a C ABI).

The compute path is libsrhip.so (csrc/), loaded through ctypes by api.py.
Importing the package never falls back to a CPU evaluator: when the shared
library is missing, the first call raises.
"""
from .api import (  # noqa: F401
    Context, Population, SrError, EVAL_PLAIN, EVAL_BATCHED, EVAL_CACHED, EVAL_JIT,
)

__all__ = ["Context", "Population", "SrError", "EVAL_PLAIN", "EVAL_BATCHED", "EVAL_CACHED",
           "EVAL_JIT"]
