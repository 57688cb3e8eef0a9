"""mp_probe_shot_reduce_noalloc (fixed library scratch, no allocation) against
mp_probe_shot_reduce (library-allocated partials) through the C-ABI: the same kernels in
the same partition, so the outputs must be bit-identical, on two calls in a row, the
second with smaller planes and fewer bins than the first."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _planes(rng, n1, n2, nb):
    from mphip import abi
    from mphip.state import DeviceBuffer
    arr = (abi.Plane * 2)()
    bufs = []
    for p in range(2):
        b = {k: DeviceBuffer.from_array(rng.standard_normal(2 * n1 * n2 * nb) * 1e-3) for k in ("ea", "eb", "ha", "hb")}
        b["energy"] = DeviceBuffer.from_array(rng.standard_normal(n1 * n2) * 1e-3)
        arr[p].axis, arr[p].index = 0, p
        arr[p].dft_ea, arr[p].dft_eb, arr[p].dft_ha, arr[p].dft_hb = (b[k].ptr for k in ("ea", "eb", "ha", "hb"))
        arr[p].energy = b["energy"].ptr
        bufs.append(b)
    return arr, bufs


def test_noalloc_shot_reduce_matches_allocating_entry_point():
    from mphip import _lib, abi
    from mphip.state import DeviceBuffer
    _lib.init(0)
    L = _lib.lib()
    rng = np.random.default_rng(7)
    for n1, n2, nb in ((37, 29, 5), (16, 12, 3)):
        print(f"case shot reduce no-alloc {n1}x{n2} nb={nb}", flush=True)
        arr, bufs = _planes(rng, n1, n2, nb)
        area = DeviceBuffer.from_array(rng.uniform(0.5, 1.5, n1 * n2))
        mode_w = DeviceBuffer.from_array(rng.standard_normal(n1 * n2))
        nout = 4 * nb + 10
        ref_out = DeviceBuffer.from_array(np.full(nout, np.nan))
        got_out = DeviceBuffer.from_array(np.full(nout, np.nan))
        q0 = nb // 2
        _lib.call("probe_shot_reduce", C.byref(arr[0]), C.byref(arr[1]), n1, n2, nb, q0, mode_w.vp, area.vp,
                  ref_out.vp)
        _lib.call("probe_shot_reduce_noalloc", C.byref(arr[0]), C.byref(arr[1]), n1, n2, nb, q0, mode_w.vp, area.vp,
                  got_out.vp)
        _lib.call("synchronize")
        want, got = ref_out.download(np.float64), got_out.download(np.float64)
        assert np.all(np.isfinite(want))
        assert np.array_equal(got, want), np.abs(got - want).max()
        for b in bufs:
            for v in b.values():
                v.free()
        for v in (area, mode_w, ref_out, got_out):
            v.free()
    # more bins than the fixed scratch holds: rejected before any device work
    arr, bufs = _planes(rng, 4, 4, 1)
    nb = abi.MP_SHOT_NOALLOC_MAX_BINS + 1
    rc = L.mp_probe_shot_reduce_noalloc(C.byref(arr[0]), C.byref(arr[1]), 4, 4, nb, 0, arr[0].energy, arr[0].energy,
                                        arr[1].energy)
    assert rc == -1   # MP_ERR_ARG
    assert b"exceeds the maximum" in L.mp_last_error()
    for b in bufs:
        for v in b.values():
            v.free()
