"""GPU parity of lane decoding (tuning lane_max_points): the RLGR of short
slices decoded one slice per lane (lane_rlgr_kernel) on the decoder's second
stream while the scalar decoder takes the long slices.

* crafted RLGR streams (runs with large exponents, escapes, mode changes)
  with every slice on the lanes: colours equal to the oracle's, bit for bit;
* random colour bytes over real trees: per-slice statuses and the outputs of
  the slices that decode equal those of the scalar decoder;
* a c1 GoP (10-bit, ~816k points/frame, GoP 8, qp 4) at several thresholds
  against the oracle, a c2 GoP (16 frames, 15 of them predictive and deferred
  to the finish pass) and 8 frames of c4 (qp 1, slices up to 4 000 points on
  the lanes) against the scalar decoder, and a batch of two GoPs (two
  segments), all in one decoder whose lane buffers grow and are reused.
"""
import sys

import numpy as np
import pytest

from _frames import pack_frame
from test_gpu_rlgr_stress import IDENT, _stream

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dims(frames):
    metas = [f.meta() for f in frames]
    return max(int(m["index"].max()) for m in metas) + 1, max(int(m["size"].max()) for m in metas)


def _decode(pvvc, dec, frames, lane_max_points):
    dec.set_tuning(lane_max_points=lane_max_points)
    b = dec.batch(frames)
    prev = pvvc.set_timing(True)  # the launch count comes from the timing events
    try:
        xyz, rgb = dec.decode_host(b, pvvc.MODE_PERSISTENT, pvvc.OUT_NATIVE)
        n_launch = len(pvvc.last_decode_kernel_times())
    finally:
        pvvc.set_timing(prev)
    st = b.status()
    offs = list(zip(b.frame_offset, b.frame_points))
    b.free()
    return xyz, rgb, st, offs, n_launch


def test_lane_crafted_streams_match_oracle(pvvc, oracle, tmp_path):
    img, syms = _stream(pvvc, oracle)
    path = tmp_path / "stress_0.frame"
    path.write_bytes(img)
    ref = oracle.StreamDecoder(fp64=False).decode(oracle.OracleFrame(str(path)), 8)
    assert (ref["status"] == 0).all()
    f = pvvc.Frame.load(str(path))
    dec = pvvc.Decoder(*_dims([f]))
    for lmax in (1 << 30, 1500):
        _say("crafted streams, lane_max_points", lmax)
        dec.set_tuning(lane_max_points=lmax)
        b = dec.batch([f])
        pts = dec.decode_host(b, pvvc.MODE_PERSISTENT, pvvc.OUT_POINTS)
        assert (b.status() == 0).all(), [syms[s][0] for s in np.flatnonzero(b.status())]
        np.testing.assert_array_equal(np.stack([pts["r"], pts["g"], pts["b"]], 1), ref["rgbf"] / np.float32(255))
        b.free()
    dec.close()


def test_lane_random_colour_bytes_equal_scalar(pvvc):
    rng = np.random.default_rng(21)
    slices = []
    for s in range(256):
        v = np.unique(rng.integers(0, 32, (int(rng.integers(2, 400)), 3)), axis=0).astype(np.float32)
        geo, leaf, _ = pvvc.encode_colors_cpu(v, np.zeros((len(v), 3), np.float32), 1)
        nbytes = int(rng.integers(0, 3 * len(leaf)))
        col = np.packbits(rng.random(8 * nbytes) < rng.choice([0.02, 0.1, 0.5, 0.9])).tobytes()
        slices.append(dict(index=s, type=0x01, size=len(leaf), qp=int(rng.integers(1, 9)), mv=IDENT, geo=geo, col=col))
    f = pvvc.Frame.parse(pack_frame(0, slices))
    dec = pvvc.Decoder(*_dims([f]))
    xa, ra, sa, _, _ = _decode(pvvc, dec, [f], -1)
    xb, rb, sb, _, _ = _decode(pvvc, dec, [f], 1 << 30)
    _say("random colour bytes: statuses", dict(zip(*np.unique(sa, return_counts=True))))
    assert len(np.unique(sa)) >= 2
    np.testing.assert_array_equal(sb, sa)
    sizes = f.meta()["size"].astype(np.int64)
    ok = np.repeat(sa == 0, sizes)
    np.testing.assert_array_equal(xb[ok], xa[ok])
    np.testing.assert_array_equal(rb[ok], ra[ok])
    dec.close()


def test_lane_c1_c2_and_two_segments(pvvc, oracle, tmp_path):
    _say("c1 GoP (seed 11), c2 GoP (seed 5, 16 frames)")
    c1 = pvvc.synth_stream("c1", n_frames=8, seed=11, threads=16)
    c2 = pvvc.synth_stream("c2", n_frames=16, seed=5, threads=16)
    f1, f2 = [pvvc.Frame.parse(b) for b in c1], [pvvc.Frame.parse(b) for b in c2]
    od = oracle.StreamDecoder(fp64=False)
    refs = []
    for k, img in enumerate(c1):
        p = tmp_path / f"c1_{k}.frame"
        p.write_bytes(img)
        refs.append(od.decode(oracle.OracleFrame(str(p)), 16))
        assert (refs[-1]["status"] == 0).all()
    dims = [max(a, b) for a, b in zip(_dims(f1), _dims(f2))]
    dec = pvvc.Decoder(*dims)
    for lmax in (1000, 3000, 1 << 30):
        xyz, rgb, st, offs, n_launch = _decode(pvvc, dec, f1, lmax)
        _say(f"  c1 lane_max_points {lmax}: statuses ok {(st == 0).all()}")
        assert (st == 0).all() and n_launch == 1
        for k, ref in enumerate(refs):
            o, n = offs[k]
            np.testing.assert_array_equal(xyz[o:o + n], ref["xyz"], err_msg=f"c1 frame {k}")
            np.testing.assert_array_equal(rgb[o:o + n], ref["rgb"], err_msg=f"c1 frame {k}")
    x0, r0, s0, _, _ = _decode(pvvc, dec, f2, -1)
    assert (s0 == 0).all()
    for lmax in (2000, 1 << 30):
        x1, r1, s1, _, n_launch = _decode(pvvc, dec, f2, lmax)
        _say(f"  c2 lane_max_points {lmax}: equal to the scalar decode")
        assert (s1 == 0).all() and n_launch == 1
        np.testing.assert_array_equal(x1, x0)
        np.testing.assert_array_equal(r1, r0)
    c4 = [pvvc.Frame.parse(b) for b in pvvc.synth_stream("c4", n_frames=8, seed=6, threads=16)]
    d4 = pvvc.Decoder(*_dims(c4))
    x0, r0, s0, _, _ = _decode(pvvc, d4, c4, -1)
    assert (s0 == 0).all()
    x1, r1, s1, _, n_launch = _decode(pvvc, d4, c4, 4000)
    _say("  c4 (8 frames, qp 1) lane_max_points 4000: equal to the scalar decode")
    assert (s1 == 0).all() and n_launch == 1
    np.testing.assert_array_equal(x1, x0)
    np.testing.assert_array_equal(r1, r0)
    d4.close()
    two = f1 + f1  # the second intra frame starts a second segment
    xa, ra, sa, _, na = _decode(pvvc, dec, two, -1)
    xb, rb, sb, _, nb = _decode(pvvc, dec, two, 1000)
    _say(f"  two c1 GoPs: {na} segments")
    assert na == nb == 2 and (sa == 0).all() and (sb == 0).all()
    np.testing.assert_array_equal(xb, xa)
    np.testing.assert_array_equal(rb, ra)
    dec.close()
