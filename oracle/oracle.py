"""ctypes loader for the CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product. The oracle exports the
core ABI under the prefix ``oracle_``, so the product's host facade can
drive it through ``d3force_amd.Core``.
"""

import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "liboracle.so")


def build():
    subprocess.check_call(["sh", os.path.join(HERE, "build.sh")])


def load():
    """Returns a d3force_amd.Core over the oracle library, building it first
    when it is missing."""
    from d3force_amd.core import Core
    if not os.path.exists(LIBRARY):
        build()
    core = Core(LIBRARY, prefix="oracle_")
    core.counters = core.bind("counters", None, [ctypes.c_int, ctypes.c_void_p])
    core.set_threads = core.bind("set_threads", None, [ctypes.c_int])
    core.fmm_stats = core.bind("fmm_stats", None, [ctypes.c_uint32, ctypes.c_void_p])
    core.tree_keys = core.bind("tree_keys", ctypes.c_void_p, [ctypes.c_uint32])
    return core
