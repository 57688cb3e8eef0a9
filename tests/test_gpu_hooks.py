"""The device side of the backend hooks a Julia HIP extension needs (src/core/Backend.jl:
49-83, src/sim/ProbeReadout.jl:291, src/fdtd/Kernels.jl:467-478) through the C-ABI.

reduce_sum / reduce_mean / reduce_count_negative / reduce_norm_mean are Julia Base
reductions in the reference; numpy's FP64 sums are their reference here (the device sums
in a fixed partition, so they agree to FP64 reassociation and repeat bit for bit).
plane_to_host is a view copy: bit for bit.  The trace samples are per-cell products: bit
for bit with numpy, and their plane sums equal mp_probe_trace_plane (checked against the
oracle elsewhere) to reassociation.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(a):
    from mphip.state import DeviceBuffer
    return DeviceBuffer.from_array(np.ascontiguousarray(a))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [0, 1, 1000, 3_000_001])
def test_reduce_matches_numpy(dtype, n):
    from mphip import _lib, abi
    _lib.init(0)
    print(f"case reduce n={n} {np.dtype(dtype).name}", flush=True)
    rng = np.random.default_rng(n + 1)
    x, y, z = (rng.standard_normal(n).astype(dtype) for _ in range(3))
    dx, dy, dz = _dev(x), _dev(y), _dev(z)
    out = _dev(np.full(1, np.nan))
    code = abi.MP_F32 if dtype == np.float32 else abi.MP_F64
    want = {abi.MP_RED_SUM: np.sum(x.astype(np.float64)),
            abi.MP_RED_COUNT_NEGATIVE: float(np.count_nonzero(x < 0)),
            abi.MP_RED_NORM_SUM: np.sum(np.sqrt(x.astype(np.float64) ** 2 + y.astype(np.float64) ** 2 +
                                                z.astype(np.float64) ** 2))}
    scale = {abi.MP_RED_SUM: np.sum(np.abs(x.astype(np.float64))), abi.MP_RED_COUNT_NEGATIVE: 0.0,
             abi.MP_RED_NORM_SUM: want[abi.MP_RED_NORM_SUM]}
    for op, w in want.items():
        _lib.call("reduce", dx.vp, dy.vp, dz.vp, code, n, op, out.vp)
        got = out.download(np.float64)[0]
        assert abs(got - w) <= 1e-12 * max(scale[op], 1.0), (op, got, w)
        _lib.call("reduce", dx.vp, dy.vp, dz.vp, code, n, op, out.vp)
        assert out.download(np.float64)[0] == got   # same partition: the same bits
    for b in (dx, dy, dz, out):
        b.free()


def test_sum_blocks_matches_numpy():
    from mphip import _lib
    _lib.init(0)
    rng = np.random.default_rng(5)
    for block, nb in ((37 * 29, 4), (1, 7), (300_000, 3), (0, 2)):
        x = rng.standard_normal(block * nb)
        dx, out = _dev(x if x.size else np.zeros(1)), _dev(np.full(nb, np.nan))
        _lib.call("sum_blocks", dx.vp, block, nb, out.vp)
        got = out.download(np.float64)
        want = x.reshape(nb, block).sum(axis=1) if block else np.zeros(nb)
        assert np.allclose(got, want, rtol=0, atol=1e-12 * max(1.0, np.abs(x).sum())), (block, nb)
        dx.free()
        out.free()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_plane_gather_matches_views(dtype):
    from mphip import _lib, abi
    _lib.init(0)
    nx, ny, nz = 13, 7, 9
    f = np.random.default_rng(2).standard_normal((nx, ny, nz)).astype(dtype)
    df = _dev(f.reshape(-1, order="F"))
    code = abi.MP_F32 if dtype == np.float32 else abi.MP_F64
    for axis, idx in ((0, 0), (0, 12), (1, 3), (1, 6), (2, 0), (2, 8)):
        view = f[(slice(None),) * axis + (idx,)]   # the reference's views: f[i,:,:], f[:,i,:], f[:,:,i]
        out = _dev(np.zeros(view.size, dtype))
        _lib.call("plane_gather", df.vp, code, nx, ny, nz, axis, idx, out.vp)
        got = out.download(dtype).reshape(view.shape, order="F")
        assert np.array_equal(got, view), (axis, idx)
        out.free()
    rc = _lib.lib().mp_plane_gather(df.vp, code, nx, ny, nz, 1, ny, df.vp)
    assert rc == -1
    df.free()


def test_trace_samples_sum_to_trace_plane():
    from mphip import _lib, abi
    from mphip.state import DeviceBuffer, FDTDState
    from test_gpu_probe import _uniform_spec
    spec = _uniform_spec(10, 23, 17)
    st = FDTDState(spec)
    rng = np.random.default_rng(9)
    ey, ez = (rng.standard_normal(spec.shape).astype(np.float32) for _ in range(2))
    st.set_field("ey", ey)
    st.set_field("ez", ez)
    ny, nz = spec.shape[1], spec.shape[2]
    w, area = rng.standard_normal(ny * nz), rng.uniform(0.5, 2.0, ny * nz)
    dw, da = _dev(w), _dev(area)
    work = DeviceBuffer.zeros(8 * 4 * ny * nz)
    _lib.call("probe_trace_samples", C.byref(st.yee), 2, 7, dw.vp, da.vp, work.vp)
    got = work.download(np.float64).reshape(4, ny * nz)
    wa = w * area
    want = np.stack([ey[2].reshape(-1, order="F").astype(np.float64) * wa,
                     ez[2].reshape(-1, order="F").astype(np.float64) * wa,
                     ey[7].reshape(-1, order="F").astype(np.float64) * wa,
                     ez[7].reshape(-1, order="F").astype(np.float64) * wa])
    assert np.array_equal(got, want)
    sums = DeviceBuffer.zeros(8 * 4)
    _lib.call("sum_blocks", work.vp, ny * nz, 4, sums.vp)
    fused = DeviceBuffer.zeros(8 * 4)
    _lib.call("probe_trace_plane", C.byref(st.yee), 0, 2, 7, dw.vp, da.vp, fused.vp)
    a, b = sums.download(np.float64), fused.download(np.float64)
    assert np.allclose(a, b, rtol=0, atol=1e-12 * np.abs(want).sum(axis=1).max())
    for x in (dw, da, work, sums, fused):
        x.free()
    st.free()
