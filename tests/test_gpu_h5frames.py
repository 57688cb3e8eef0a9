"""Movie frames of the HDF5 writers staged on the device (mp_plane_abs, mp_probe_frames):
each frame in the file equals |field| of the downloaded volume on its plane, as float32,
bit for bit (the reference's Float32.(abs.(view)), ext/MagnetoPhotonicHDF5Ext.jl:619-697)."""
import numpy as np
import pytest

from mphip import configs, h5out

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not h5out.available(), reason="no HDF5 C library in this image")]


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_frame_writers_stage_on_device(tmp_path, dtype):
    from mphip.state import FDTDState
    spec, meta = configs.small_coupled(dtype=dtype, steps=60, amplitude=2e8, n=24)
    print(f"case HDF5 movie frames staged on the device, {meta['name']} {dtype}", flush=True)
    gpu = FDTDState(spec)
    g = spec.grid
    xs = [float(c) * 1e6 for c in g.x.centers[[3, 9, 14, 20]]]
    p = str(tmp_path / "frames.h5")
    w = h5out.production_frame_writer(p, gpu, frame_count=3, slice_z_pos=float(g.z.centers[7]),
                                      slice_y_pos=float(g.y.centers[11]), x_um=xs)
    pw = h5out.probe_frame_writer(p, gpu, group="probe/initial", frame_count=3, pre_x=5, post_x=18)
    expect, pexpect = [], []
    for _ in range(3):
        gpu.run(7)
        w.record(gpu)
        pw.record(gpu)
        gpu.synchronize()
        ey, ez = np.abs(gpu.field("ey")), np.abs(gpu.field("ez"))
        expect.append({"Ez_xy": ez[:, :, w.kz], "Ez_xz": ez[:, w.jy, :],
                       **{k: ez[i, :, :] for k, i in zip(w.KEYS[2:], w.ix)}})
        pexpect.append({"Ey_xy": ey[:, :, pw.kz], "Ez_xy": ez[:, :, pw.kz], "Ey_yz_pre": ey[pw.ipre, :, :],
                        "Ez_yz_pre": ez[pw.ipre, :, :], "Ey_yz_post": ey[pw.ipost, :, :],
                        "Ez_yz_post": ez[pw.ipost, :, :]})
    w.record(gpu)                                    # past frame_count: ignored
    w.close()
    pw.close()
    assert w.kz == 7 and w.jy == 11 and w.ix == (3, 9, 14, 20)
    with h5out.H5File(p, "r") as f:
        for k in w.KEYS:
            cube = f.read(f"pump/{k}")
            for i in range(3):
                assert np.array_equal(cube[:, :, i], expect[i][k].astype(np.float32)), (k, i)
        for k in pw.KEYS:
            cube = f.read(f"probe/initial/{k}")
            for i in range(3):
                assert np.array_equal(cube[:, :, i], pexpect[i][k].astype(np.float32)), (k, i)
    assert max(float(expect[-1]["Ez_xy"].max()), float(pexpect[-1]["Ey_xy"].max())) > 0.0
    gpu.free()
