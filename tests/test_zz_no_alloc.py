"""The allocation-free dot entry point (flux_dot_noalloc) against the
reducer it mirrors (flux_dot_create + flux_dot_read), through the C-ABI:
two calls in a row, the second on a smaller input, bit for bit; the
scratch limit is enforced with an error; run alone or after other tests."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _reference(F, shape, a, b):
    red = F.create_dot_reducer(shape, a, b)
    return red.read(), red.partial.to_array()


def _noalloc(F, shape, a, b, lanes):
    lib = F.lib()
    lanes_out = F.array_n("float", (lanes,))
    total = F.array_n("f64", (1,))
    sh = (C.c_int * len(shape))(*shape)
    rc = lib.flux_dot_noalloc(len(shape), sh, C.c_void_p(a.h), C.c_void_p(b.h), C.c_void_p(lanes_out.h),
                              C.c_void_p(total.h))
    assert rc == 0, lib.flux_last_error().decode()
    F.sync()
    return float(total.to_array()[0]), lanes_out.to_array()


def test_noalloc_matches_reducer_twice(flux):
    F = flux
    rng = np.random.default_rng(77)
    for shape in ((256, 128, 64), (96, 40, 8)):  # the second call is smaller
        n = int(np.prod(shape))
        a = F.array_n("float", shape, rng.standard_normal(n).astype(np.float32))
        b = F.array_n("float", shape, rng.standard_normal(n).astype(np.float32))
        want_t, want_l = _reference(F, shape, a, b)
        got_t, got_l = _noalloc(F, shape, a, b, shape[0])
        print(shape, got_t, want_t)
        assert got_t == want_t
        np.testing.assert_array_equal(got_l, want_l)


def test_noalloc_probes_and_limit(flux):
    F = flux
    shape = (6, 4, 5)
    ones = F.array_n("float", shape).fill(1.0)
    ramp = F.array_n("float", shape, np.arange(120, dtype=np.float32))
    t, lanes = _noalloc(F, shape, ramp, ones, 6)
    assert t == 7140.0
    np.testing.assert_array_equal(lanes, np.arange(120).reshape(20, 6).sum(axis=0).astype(np.float32))
    big = (140000, 2)
    a = F.array_n("float", big)
    lib = F.lib()
    total = F.array_n("f64", (1,))
    sh = (C.c_int * 2)(*big)
    rc = lib.flux_dot_noalloc(2, sh, C.c_void_p(a.h), C.c_void_p(a.h), None, C.c_void_p(total.h))
    assert rc != 0 and "scratch" in lib.flux_last_error().decode()
