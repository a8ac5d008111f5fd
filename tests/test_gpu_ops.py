"""The reference's per-particle operators (batched kernels behind the C ABI)
against the oracle's per-particle restatements, on random inputs."""
import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(77)


@pytest.fixture(scope="module")
def L():
    from ohhelppic3d_amd import _lib
    if not _lib.available():
        pytest.fail("libohpic.so is not built: the GPU path has no fallback")
    return _lib


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def test_boris_batch_matches_oracle(L, orc):
    n = 500
    x, y, z = (RNG.uniform(0, 8, n) for _ in range(3))
    v = RNG.standard_normal((3, n))
    eb = RNG.standard_normal((n, 6))
    arr = [np.array(a, copy=True) for a in (x, y, z, v[0], v[1], v[2])]
    rc = L.lib().ohpic_boris_batch(0, n, *[L.dptr(a) for a in arr],
                                   L.dptr(np.ascontiguousarray(eb.reshape(-1))), -0.7, 0.1)
    L.check(rc)
    for i in range(n):
        pos, vel = orc.boris_move([x[i], y[i], z[i]], v[:, i], -0.7, eb[i], 0.1)
        assert np.allclose([arr[0][i], arr[1][i], arr[2][i]], pos, rtol=1e-14, atol=1e-14)
        assert np.allclose([arr[3][i], arr[4][i], arr[5][i]], vel, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("wo", [0, 1])
def test_interp_and_scatter_batch_match_oracle(L, orc, wo):
    nelem, lb, ub = 6, [-1, -1, -1], [9, 7, 5]
    shape = (nelem, 11, 9, 7)
    values = RNG.standard_normal(shape)
    flat = np.ascontiguousarray(values.reshape(-1, order="F"))
    n = 300
    pos = np.stack([RNG.uniform(-1, 8, n), RNG.uniform(-1, 6, n), RNG.uniform(-1, 4, n)], 1)
    out = np.zeros(n * nelem)
    L.check(L.lib().ohpic_interp_batch(0, n, L.dptr(np.ascontiguousarray(pos.reshape(-1))),
                                       L.dptr(flat), nelem, L.i32ptr(_i32(lb)),
                                       L.i32ptr(_i32(ub)), wo, L.dptr(out)))
    out = out.reshape(n, nelem)
    for i in range(n):
        assert np.allclose(out[i], orc.interp(values, lb, ub, pos[i], wo), rtol=1e-13,
                           atol=1e-14)
    amount = np.array([0.5, -1.0, 2.0, 0.0, 1.5, -0.25])
    acc = np.zeros_like(flat)
    L.check(L.lib().ohpic_scatter_batch(0, n, L.dptr(np.ascontiguousarray(pos.reshape(-1))),
                                        L.dptr(acc), nelem, L.i32ptr(_i32(lb)),
                                        L.i32ptr(_i32(ub)), L.dptr(amount), wo))
    ref = np.zeros_like(flat)
    for i in range(n):
        orc.lib().orc_scatter(ref, nelem, _i32(lb), _i32(ub), pos[i], amount, wo)
    assert np.allclose(acc, ref, rtol=1e-13, atol=1e-13)


def test_stencil_outside_bounds_is_an_error(L):
    values = np.zeros(1 * 4 * 4 * 4)
    pos = np.array([3.5, 1.0, 1.0])  # upper node 4 > ub 3
    out = np.zeros(1)
    rc = L.lib().ohpic_interp_batch(0, 1, L.dptr(pos), L.dptr(values), 1,
                                    L.i32ptr(_i32([0, 0, 0])), L.i32ptr(_i32([3, 3, 3])), 0,
                                    L.dptr(out))
    assert rc == L.EINVAL
    assert b"outside" in L.lib().ohpic_last_error(None)


def test_wrap_batch_special_values_bit_exact(L, orc):
    n = 16
    p = np.array([3.25, -0.5, 16.0, 16.75, -5e-324, np.nextafter(16.0, 0.0), -0.0, -15.75,
                  -20.0, 40.0])
    for apps in (1, 2):
        got = p.copy()
        L.check(L.lib().ohpic_wrap_batch(0, len(p), L.dptr(got), n, apps))
        for a, b in zip(got, p):
            r = b
            for _ in range(apps):
                r = orc.apply_periodic(n, r)
            assert a == r and math.copysign(1, a) == math.copysign(1, r)
