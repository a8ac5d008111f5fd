"""The allocation-free transform ohpic_fft3d_r2r_noalloc, through the C ABI,
against the allocating ohpic_fft3d_r2r: bit for bit equal on consecutive
calls (each followed by a smaller one, which must not see the previous
call's data), forward and backward, up to the maximum size; sizes beyond the
maximum are refused before any device work."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


@pytest.fixture(scope="module")
def L():
    from ohhelppic3d_amd import _lib
    if not _lib.available():
        pytest.fail("libohpic.so is not built: the GPU path has no fallback")
    return _lib


def _both(L, n, backward, a):
    lib = L.lib()
    ng = n[0] * n[1] * n[2]
    ref = np.full(ng, np.nan)
    got = np.full(ng, np.nan)
    assert lib.ohpic_fft3d_r2r(0, *n, backward, L.dptr(a), L.dptr(ref)) == L.OK
    assert lib.ohpic_fft3d_r2r_noalloc(0, *n, backward, L.dptr(a), L.dptr(got)) == L.OK, \
        lib.ohpic_last_error(None)
    return ref, got


def test_oversize_and_bad_arguments_are_refused(L):
    lib = L.lib()
    a = np.zeros(8)
    # 129 * 128 * 128 > 2^21 points: refused with the limit in the message
    assert lib.ohpic_fft3d_r2r_noalloc(0, 129, 128, 128, 0, L.dptr(a), L.dptr(a)) == L.EINVAL
    assert b"allocation-free maximum" in lib.ohpic_last_error(None)
    assert lib.ohpic_fft3d_r2r_noalloc(0, 2, 2, 2, 0, None, L.dptr(a)) == L.EINVAL
    assert lib.ohpic_fft3d_r2r_noalloc(0, 0, 2, 2, 0, L.dptr(a), L.dptr(a)) == L.EINVAL
    # an axis with a prime factor above 13
    b = np.zeros(17 * 2 * 2)
    assert lib.ohpic_fft3d_r2r_noalloc(0, 17, 2, 2, 0, L.dptr(b), L.dptr(b)) == L.EINVAL


@pytest.mark.gpu
@pytest.mark.parametrize("backward", [0, 1])
def test_noalloc_equals_allocating_transform(L, backward):
    rng = np.random.default_rng(11 + backward)
    # each call followed by a smaller one; radix 2/3/4/5 axes and the maximum
    for n in ((64, 48, 40), (24, 16, 10), (128, 128, 128), (30, 12, 7), (8, 8, 8)):
        a = rng.standard_normal(n[0] * n[1] * n[2])
        ref, got = _both(L, n, backward, a)
        print(n, "backward" if backward else "forward", "max |diff|",
              float(np.max(np.abs(got - ref))))
        assert np.array_equal(got, ref), n


@pytest.mark.gpu
def test_noalloc_round_trip_and_in_place(L):
    lib = L.lib()
    n = (96, 64, 32)
    a = np.random.default_rng(3).standard_normal(n[0] * n[1] * n[2])
    buf = a.copy()
    # in and out may alias
    assert lib.ohpic_fft3d_r2r_noalloc(0, *n, 0, L.dptr(buf), L.dptr(buf)) == L.OK
    assert lib.ohpic_fft3d_r2r_noalloc(0, *n, 1, L.dptr(buf), L.dptr(buf)) == L.OK
    assert np.max(np.abs(buf - a)) <= 1e-13 * np.max(np.abs(a))
