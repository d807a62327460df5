"""GPU parity at full configuration shapes, whole GoPs per launch.

* c1 (10-bit, ~816k points/frame, ~860 patches of 200-3000 points, GoP 8,
  qp 4): one whole seeded GoP through every decode form -- the batched
  persistent launch, per-frame launches, the captured hipGraph, and the
  drop-in launch_cudaProcess -- against the CPU oracle.  Geometry bit-exact,
  RGB8 exact (and |diff| <= 1 asserted), drop-in colours equal to the
  oracle's FP32 RGB / 255.  The FP64 inverse-RAHT error (max, RMS) and the
  exact-match fraction are printed.
* c2 (11-bit, ~3.3M points/frame, patches 50-20000, GoP 16): the whole
  16-frame GoP in ONE persistent launch (15 predictive frames deferred to
  the finish pass).
* c4 (12-bit, 8 humanoids, ~5.5M points/frame, qp 1, patches 20-50000,
  GoP 32): the first 8 frames of its GoP in one persistent launch.

The reference harness decodes frame by frame in decode order
(src/codec/pvvc_decompression.cc:105-149); the oracle restates it.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _write(tmp, name, imgs):
    paths = []
    for k, b in enumerate(imgs):
        p = tmp / f"{name}_{k}.frame"
        p.write_bytes(b)
        paths.append(p)
    return paths


def _dims(frames):
    metas = [f.meta() for f in frames]
    return max(int(m["index"].max()) for m in metas) + 1, max(int(m["size"].max()) for m in metas)


@pytest.fixture(scope="module")
def c1_gop(pvvc, oracle, tmp_path_factory):
    _say("c1 GoP: generating 8 frames (seed 11)")
    imgs = pvvc.synth_stream("c1", n_frames=8, seed=11, threads=16)
    paths = _write(tmp_path_factory.mktemp("c1"), "c1", imgs)
    dec = oracle.StreamDecoder(fp64=True)
    refs = [dec.decode(oracle.OracleFrame(str(p)), 16) for p in paths]
    for r in refs:
        assert (r["status"] == 0).all()
    frames = [pvvc.Frame.load(str(p)) for p in paths]
    _say("  points/frame", [f.points for f in frames], "patches/frame", frames[0].slice_cnt)
    return frames, refs


def _fp64_report(tag, rgb, refs, offsets):
    exact, cnt, diffs = 0, 0, []
    for (o, n), ref in zip(offsets, refs):
        g = rgb[o:o + n]
        exact += int((g == ref["rgb"]).all(1).sum())
        cnt += n
        diffs.append(g.astype(np.float64) - np.clip(ref["rgb64"], 0.0, 255.0))
    d = np.concatenate(diffs)
    _say(f"  {tag}: RGB8 exact vs FP32 oracle {exact / cnt:.6f} of {cnt} points; vs FP64 inverse RAHT "
         f"max |err| {np.abs(d).max():.4f}, RMS {np.sqrt((d ** 2).mean()):.4f}")
    return np.abs(d).max()


@pytest.mark.parametrize("mode", [1, 0, 2], ids=["persistent", "per_frame", "graph"])
def test_c1_gop_native_matches_oracle(pvvc, c1_gop, mode):
    frames, refs = c1_gop
    _say("c1 GoP native, mode", mode)
    dec = pvvc.Decoder(*_dims(frames))
    b = dec.batch(frames)
    prev = pvvc.set_timing(True)  # launch counts come from the timing events
    try:
        xyz, rgb = dec.decode_host(b, mode, pvvc.OUT_NATIVE)
        n_launch = len(pvvc.last_decode_kernel_times())
    finally:
        pvvc.set_timing(prev)
    assert (b.status() == 0).all(), np.unique(b.status(), return_counts=True)
    if mode == pvvc.MODE_PERSISTENT:
        # the whole GoP is one persistent launch: a single intra frame, no segment split
        assert n_launch == 1
    offs = list(zip(b.frame_offset, b.frame_points))
    for k, ref in enumerate(refs):
        o, n = offs[k]
        np.testing.assert_array_equal(xyz[o:o + n], ref["xyz"], err_msg=f"frame {k} geometry")
        assert np.abs(rgb[o:o + n].astype(int) - ref["rgb"].astype(int)).max() <= 1
        np.testing.assert_array_equal(rgb[o:o + n], ref["rgb"], err_msg=f"frame {k} colour")
    # the rounding to RGB8 itself is within 0.5 of the FP64 transform (+ FP32 slack)
    assert _fp64_report(f"mode {mode}", rgb, refs, offs) < 0.51


def test_c1_gop_dropin_matches_oracle(pvvc, c1_gop):
    frames, refs = c1_gop
    st = pvvc.State(*_dims(frames))
    out = pvvc.DeviceBuffer(24 * max(f.points for f in frames))
    for k, (f, ref) in enumerate(zip(frames, refs)):
        _say("c1 drop-in frame", k)
        pts = pvvc.decode_dropin(f, st, out)
        assert st.synchronize() == 0
        np.testing.assert_array_equal(np.stack([pts["x"], pts["y"], pts["z"]], 1), ref["xyz"])
        np.testing.assert_array_equal(np.stack([pts["r"], pts["g"], pts["b"]], 1), ref["rgbf"] / np.float32(255))


@pytest.mark.parametrize("preset,n_frames,seed", [("c2", 16, 5), ("c4", 8, 6)])
def test_large_gop_one_launch_matches_oracle(pvvc, oracle, tmp_path, preset, n_frames, seed):
    _say(f"{preset}: generating {n_frames} frames of one GoP (seed {seed})")
    gop = pvvc.SYNTH_PRESETS[preset]["gop"]
    assert n_frames <= gop
    imgs = pvvc.synth_stream(preset, n_frames=n_frames, seed=seed, threads=16)
    paths = _write(tmp_path, preset, imgs)
    frames = [pvvc.Frame.load(str(p)) for p in paths]
    sizes = np.concatenate([f.meta()["size"] for f in frames[:1]])
    _say(f"  points/frame {frames[0].points}, patches {frames[0].slice_cnt}, sizes {sizes.min()}-{sizes.max()}")
    dec = pvvc.Decoder(*_dims(frames))
    b = dec.batch(frames)
    prev = pvvc.set_timing(True)  # the launch count comes from the timing events
    try:
        xyz, rgb = dec.decode_host(b, pvvc.MODE_PERSISTENT, pvvc.OUT_NATIVE)
        n_launch = len(pvvc.last_decode_kernel_times())
    finally:
        pvvc.set_timing(prev)
    assert n_launch == 1, "the GoP must decode in one persistent launch"
    assert (b.status() == 0).all(), np.unique(b.status(), return_counts=True)
    od = oracle.StreamDecoder(fp64=False)
    for k, p in enumerate(paths):
        ref = od.decode(oracle.OracleFrame(str(p)), 16)
        assert (ref["status"] == 0).all()
        o, n = b.frame_offset[k], b.frame_points[k]
        np.testing.assert_array_equal(xyz[o:o + n], ref["xyz"], err_msg=f"{preset} frame {k} geometry")
        np.testing.assert_array_equal(rgb[o:o + n], ref["rgb"], err_msg=f"{preset} frame {k} colour")
        _say(f"  frame {k}: {n} points equal")
