"""The oracle's BF16 / binary16 storage conversions against independent library routines.

binary16 is checked against numpy's float32 -> float16 cast (IEEE round to nearest even,
subnormals, overflow to inf) and bfloat16 against torch's CPU cast, on values that hit
every branch: normals, ties at both parities, the subnormal range, the overflow edge,
zeros of both signs, infinities and NaN.  The scaled form (value*scale stored, stored*
(1/scale) read) is checked on the scales the reduced-precision form uses.
"""
import ctypes as C

import numpy as np
import pytest


def _lib():
    import driver
    from mphip import abi
    return driver.lib(abi)


def _store(kind, x, sc=1.0):
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty(x.size, np.uint16)
    _lib().orc_store_reduced(kind, x.ctypes.data, out.ctypes.data, x.size, sc)
    return out


def _load(kind, h, isc=1.0):
    h = np.ascontiguousarray(h, np.uint16)
    out = np.empty(h.size, np.float32)
    _lib().orc_load_reduced(kind, h.ctypes.data, out.ctypes.data, h.size, isc)
    return out


def _probe_values():
    rng = np.random.default_rng(7)
    parts = [
        rng.standard_normal(20000).astype(np.float32) * np.float32(1000.0),
        (rng.standard_normal(20000) * 10.0 ** rng.uniform(-12, 6, 20000)).astype(np.float32),
        np.array([0.0, -0.0, np.inf, -np.inf, 65504.0, 65519.99, 65520.0, 70000.0, 6.1035156e-05,
                  6.0e-05, 5.96e-08, 2.98e-08, 2.9802322e-08, 1.0e-30, 1.0 + 2.0 ** -11,
                  1.0 + 3 * 2.0 ** -11, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 3.0e38], np.float32),
    ]
    # exact ties of binary16 in the normal and subnormal ranges
    m = rng.integers(0, 1024, 2000)
    e = rng.integers(-14, 15, 2000)
    parts.append(((1024 + m + 0.5) * 2.0 ** (e - 10)).astype(np.float32))
    parts.append(((m + 0.5) * 2.0 ** -24).astype(np.float32))
    v = np.concatenate(parts)
    return np.concatenate([v, -v])


def test_binary16_matches_ieee_cast():
    v = _probe_values()
    with np.errstate(over="ignore"):
        want = v.astype(np.float16).view(np.uint16)
    got = _store(3, v)
    assert np.array_equal(got, want)
    back = _load(3, want)
    with np.errstate(over="ignore"):
        ref = want.view(np.float16).astype(np.float32)
    assert np.array_equal(back.view(np.uint32), ref.view(np.uint32))


def test_bfloat16_matches_torch_cast():
    torch = pytest.importorskip("torch")
    v = _probe_values()
    want = torch.from_numpy(v.copy()).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    got = _store(2, v)
    assert np.array_equal(got, want)
    back = _load(2, want)
    ref = torch.from_numpy(want.view(np.int16).copy()).view(torch.bfloat16).float().numpy()
    assert np.array_equal(back.view(np.uint32), ref.view(np.uint32))


def test_nan_stays_nan():
    v = np.array([np.nan, -np.nan], np.float32)
    for kind in (2, 3):
        assert np.isnan(_load(kind, _store(kind, v))).all()


@pytest.mark.parametrize("kind,sc", [(3, 2.0 ** -52), (3, 2.0 ** 13), (2, 1.0)])
def test_scaled_round_trip(kind, sc):
    rng = np.random.default_rng(11)
    mag = {2.0 ** -52: 1e19, 2.0 ** 13: 0.6, 1.0: 1e6}[sc]
    x = (rng.uniform(-1, 1, 5000) * mag).astype(np.float32)
    y = _load(kind, _store(kind, x, sc), np.float32(1.0 / sc))
    rel = 2.0 ** -11 if kind == 3 else 2.0 ** -8
    big = np.abs(x) > mag * 1e-3
    assert np.all(np.abs(y[big] - x[big]) <= rel * np.abs(x[big]))
