"""force_direct_sum_static, the entry point that allocates nothing, against
force_direct_sum (the same FP64 direct sum, with device staging allocated
per call): bit-identical on two calls in a row, the second with a smaller
input; rejects more samples than its fixed scratch holds and a force whose
parameters are not on the device yet, with a clear error."""

import ctypes

import numpy as np
import pytest

from d3force_amd import core as C
from pins import Raw

pytestmark = pytest.mark.gpu

MAX = 16384


def bind(core):
    return core.bind("force_direct_sum_static", ctypes.c_int32,
                     [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, ctypes.c_void_p, ctypes.c_uint32,
                      ctypes.c_void_p])


def make(core, n, seed):
    r = np.random.default_rng(seed)
    s = Raw(core, n)
    s.put(r.uniform(-400, 400, n), r.uniform(-400, 400, n))
    fh = s.force(C.MANY_BODY, n)
    s.f64(fh, C.F0, n)[:] = r.uniform(-40, 10, n)
    s.scalar(fh, C.DISTANCE_MIN2, 4.0)
    s.scalar(fh, C.DISTANCE_MAX2, np.inf)
    return s, fh


def both(core, static, s, fh, sample, alpha):
    k = len(sample)
    sample = np.ascontiguousarray(sample, dtype=np.uint32)
    a = np.full(2 * k, np.nan)
    b = np.full(2 * k, np.nan)
    assert core.force_direct_sum(s.h, fh, alpha, sample.ctypes.data, k, a.ctypes.data) == 0
    assert static(s.h, fh, alpha, sample.ctypes.data, k, b.ctypes.data) == 0, core.last_error()
    return a, b


def test_no_alloc_direct_sum_matches_the_allocating_one(gpu_core):
    print("no-alloc direct sum: two calls", flush=True)
    static = bind(gpu_core)
    r = np.random.default_rng(1)
    # first call: 60000 nodes, the maximum number of samples
    s, fh = make(gpu_core, 60000, 2)
    a, b = both(gpu_core, static, s, fh, r.choice(60000, MAX, replace=False), 0.7)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    assert np.all(np.isfinite(b))
    s.close()
    # second call, smaller: 5000 nodes, 300 samples (the scratch keeps the
    # first call's values beyond them; they must not leak into the result)
    s, fh = make(gpu_core, 5000, 3)
    a, b = both(gpu_core, static, s, fh, r.choice(5000, 300, replace=False), 0.3)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    s.close()


def test_no_alloc_direct_sum_refuses_what_it_cannot_hold(gpu_core):
    print("no-alloc direct sum: refusals", flush=True)
    static = bind(gpu_core)
    s, fh = make(gpu_core, 20000, 4)
    big = np.arange(MAX + 1, dtype=np.uint32)
    out = np.zeros(2 * (MAX + 1))
    assert static(s.h, fh, 0.5, big.ctypes.data, MAX + 1, out.ctypes.data) == C.BAD_ARGUMENT
    assert b"exceed its fixed scratch" in gpu_core.last_error()
    s.close()
    # a fresh force whose strengths never reached the device
    s, fh = make(gpu_core, 1000, 5)
    few = np.arange(10, dtype=np.uint32)
    out = np.zeros(20)
    assert static(s.h, fh, 0.5, few.ctypes.data, 10, out.ctypes.data) == C.BAD_ARGUMENT
    assert b"not on the device" in gpu_core.last_error()
    # after the allocating call put them there, it runs
    ref = np.zeros(20)
    assert gpu_core.force_direct_sum(s.h, fh, 0.5, few.ctypes.data, 10, ref.ctypes.data) == 0
    assert static(s.h, fh, 0.5, few.ctypes.data, 10, out.ctypes.data) == 0
    assert np.array_equal(ref.view(np.uint64), out.view(np.uint64))
    s.close()
