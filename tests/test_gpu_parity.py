"""GPU parity: the HIP decoder against the CPU oracle on the same streams.

Bar: geometry (xyz after motion) bit-exact as float values; colours of the
native path bit-exact RGB8 (the stated bar is |diff| <= 1, asserted too);
drop-in Points colours equal to the oracle's unrounded FP32 RGB / 255.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _oracle_decode(oracle, paths, threads=8):
    dec = oracle.StreamDecoder(fp64=True)
    outs = []
    for p in paths:
        f = oracle.OracleFrame(str(p))
        assert f.rc == 0
        outs.append(dec.decode(f, threads))
    return outs


@pytest.fixture(scope="module")
def c0_oracle(oracle, c0_files):
    return _oracle_decode(oracle, c0_files)


@pytest.mark.parametrize("waves", [2, 1, 4])
def test_dropin_c0_matches_oracle(pvvc, c0_files, c0_oracle, waves):
    _say("dropin c0: frames", len(c0_files), "waves", waves)
    assert pvvc.hip_lib().pvvc_set_dropin_tuning(waves, 32 * 1024) == 0
    frames = [pvvc.Frame.load(str(p)) for p in c0_files]
    n_patches = max(int(f.meta()["index"].max()) for f in frames) + 1
    max_pts = max(int(f.meta()["size"].max()) for f in frames)
    st = pvvc.State(n_patches, max_pts)
    for i, (f, ref) in enumerate(zip(frames, c0_oracle)):
        _say("  frame", i, "points", f.points)
        pts = pvvc.decode_dropin(f, st)
        assert (ref["status"] == 0).all()
        for s in range(f.slice_cnt):
            assert st.status(int(f.meta()["index"][s])) == 0
        xyz = np.stack([pts["x"], pts["y"], pts["z"]], 1)
        np.testing.assert_array_equal(xyz, ref["xyz"])
        rgbf = np.stack([pts["r"], pts["g"], pts["b"]], 1)
        np.testing.assert_array_equal(rgbf, ref["rgbf"] / np.float32(255))
    assert pvvc.hip_lib().pvvc_set_dropin_tuning(4, 32 * 1024) == 0  # the default


@pytest.mark.parametrize("mode", [0, 1, 2], ids=["per_frame", "persistent", "graph"])
def test_native_c0_matches_oracle(pvvc, c0_files, c0_oracle, mode):
    _say("native c0 mode", mode)
    frames = [pvvc.Frame.load(str(p)) for p in c0_files]
    n_patches = max(int(f.meta()["index"].max()) for f in frames) + 1
    max_pts = max(int(f.meta()["size"].max()) for f in frames)
    dec = pvvc.Decoder(n_patches, max_pts)
    b = dec.batch(frames)
    xyz, rgb = dec.decode_host(b, mode, pvvc.OUT_NATIVE)
    assert (b.status() == 0).all(), np.unique(b.status())
    for i, ref in enumerate(c0_oracle):
        o, n = b.frame_offset[i], b.frame_points[i]
        np.testing.assert_array_equal(xyz[o:o + n], ref["xyz"])
        d = np.abs(rgb[o:o + n].astype(int) - ref["rgb"].astype(int))
        assert d.max() <= 1
        np.testing.assert_array_equal(rgb[o:o + n], ref["rgb"])


def test_timing_events_opt_in(pvvc, c0_files, c0_oracle):
    """Kernel timing events are recorded only while timing is on for the
    calling thread (off by default); the output is the oracle's either way."""
    _say("timing on/off, persistent c0")
    frames = [pvvc.Frame.load(str(p)) for p in c0_files]
    n_patches = max(int(f.meta()["index"].max()) for f in frames) + 1
    max_pts = max(int(f.meta()["size"].max()) for f in frames)
    dec = pvvc.Decoder(n_patches, max_pts)
    b = dec.batch(frames)
    assert pvvc.set_timing(False) is False  # the default
    outs = []
    for on in (False, True, False):
        pvvc.set_timing(on)
        xyz, rgb = dec.decode_host(b, pvvc.MODE_PERSISTENT, pvvc.OUT_NATIVE)
        kt = pvvc.last_decode_kernel_times()
        if on:
            assert len(kt) >= 1 and all(t > 0 for t in kt) and pvvc.last_kernel_ms() >= sum(kt)
        else:
            assert kt == [] and pvvc.last_kernel_ms() == 0
        outs.append((xyz, rgb))
    pvvc.set_timing(False)
    for xyz, rgb in outs:
        for i, ref in enumerate(c0_oracle):
            o, n = b.frame_offset[i], b.frame_points[i]
            np.testing.assert_array_equal(xyz[o:o + n], ref["xyz"])
            np.testing.assert_array_equal(rgb[o:o + n], ref["rgb"])


@pytest.mark.parametrize("waves,lds", [(1, 4096), (2, 48 * 1024), (4, 16384)])
def test_tuning_variants_identical(pvvc, c0_files, c0_oracle, waves, lds):
    """waves-per-patch and LDS budgets (small budgets spill to the global
    workspace) give identical output."""
    _say("tuning", waves, lds)
    frames = [pvvc.Frame.load(str(p)) for p in c0_files]
    n_patches = max(int(f.meta()["index"].max()) for f in frames) + 1
    max_pts = max(int(f.meta()["size"].max()) for f in frames)
    dec = pvvc.Decoder(n_patches, max_pts, tuning=dict(waves_per_patch=waves, lds_bytes=lds))
    b = dec.batch(frames)
    xyz, rgb = dec.decode_host(b, pvvc.MODE_PERSISTENT, pvvc.OUT_NATIVE)
    assert (b.status() == 0).all()
    for i, ref in enumerate(c0_oracle):
        o, n = b.frame_offset[i], b.frame_points[i]
        np.testing.assert_array_equal(xyz[o:o + n], ref["xyz"])
        np.testing.assert_array_equal(rgb[o:o + n], ref["rgb"])


def test_autotune_keeps_output(pvvc, c0_files, c0_oracle):
    """The autotuner returns a configuration from its search space, sets it,
    and the decode under it still equals the oracle."""
    _say("autotune c0")
    frames = [pvvc.Frame.load(str(p)) for p in c0_files]
    n_patches = max(int(f.meta()["index"].max()) for f in frames) + 1
    max_pts = max(int(f.meta()["size"].max()) for f in frames)
    dec = pvvc.Decoder(n_patches, max_pts)
    b = dec.batch(frames)
    out = dec.outputs(b, pvvc.OUT_NATIVE)
    best, ms = dec.autotune(b, pvvc.MODE_PERSISTENT, out, reps=2)
    _say("  best", best, "ms", ms)
    assert best == dec.tuning()
    assert best["waves_per_patch"] in (1, 2, 4)
    assert 4096 <= best["lds_bytes"] <= 65536
    assert best["lane_max_points"] in (0, 256, 1000, 2000, 4000)
    assert ms > 0
    n = b.total_points
    xyz = out.xyz.read(np.float32, 3 * n).reshape(n, 3)
    rgb = out.rgb.read(np.uint8, 3 * n).reshape(n, 3)
    assert (b.status() == 0).all()
    for i, ref in enumerate(c0_oracle):
        o, k = b.frame_offset[i], b.frame_points[i]
        np.testing.assert_array_equal(xyz[o:o + k], ref["xyz"])
        np.testing.assert_array_equal(rgb[o:o + k], ref["rgb"])


def test_dropin_stage_to_matches_stage(pvvc, c0_files, c0_oracle):
    """Staging into reusable caller buffers (pinned host + device, no
    allocation per frame) decodes exactly like pvvc_dropin_stage."""
    frames = [pvvc.Frame.load(str(p)) for p in c0_files]
    n_patches = max(int(f.meta()["index"].max()) for f in frames) + 1
    max_pts = max(int(f.meta()["size"].max()) for f in frames)
    cap = max(pvvc.Stager.bytes_for(f) for f in frames)
    stager = pvvc.Stager(cap)
    st = pvvc.State(n_patches, max_pts)
    out = pvvc.DeviceBuffer(24 * max(f.points for f in frames))
    for f, ref in zip(frames, c0_oracle):
        s = stager.stage(f)
        pvvc.launch_dropin(s, st, out)
        pvvc.synchronize()
        pts = out.read(pvvc.POINTS_DTYPE, f.points)
        np.testing.assert_array_equal(np.stack([pts["x"], pts["y"], pts["z"]], 1), ref["xyz"])
        np.testing.assert_array_equal(np.stack([pts["r"], pts["g"], pts["b"]], 1), ref["rgbf"] / np.float32(255))
    with pytest.raises(pvvc.PvvcError):
        pvvc.Stager(64).stage(frames[0])  # too small: capacity error, nothing copied
    stager.close()
