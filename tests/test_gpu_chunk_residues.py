"""Slice sizes n = 1 .. 70, so that the colour stream's last 64-symbol chunk
holds every possible count 3n mod 64 = 1 .. 63 (and full chunks): the
decoder keeps a slice's last, partial chunk rotated in its accumulator lanes
and undoes the rotation when it stores the chunk.  Every decode mode and
drop-in launch shape against the oracle, bit for bit."""
import sys

import numpy as np
import pytest

from _frames import pack_frame

IDENT = [1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _frame(pvvc, rng):
    slices = []
    for k, n in enumerate(range(1, 71)):
        while True:
            v = np.unique(rng.integers(0, 16, (n + 8, 3)), axis=0)[:n].astype(np.float32)
            if len(v) == n:
                break
        yuv = (128 + rng.normal(0, 20, (n, 3))).astype(np.float32)
        geo, leaf, col = pvvc.encode_colors_cpu(v, yuv, 2)
        slices.append(dict(index=k, type=0x01, size=len(leaf), qp=2, mv=IDENT, geo=geo, col=col))
    return pack_frame(0, slices)


def test_residue_frame_covers_every_last_chunk_count(pvvc):
    import struct
    img = _frame(pvvc, np.random.default_rng(4))
    f = pvvc.Frame.parse(img)
    sizes = f.meta()["size"].astype(np.int64)
    assert set((3 * sizes) % 64) == set(range(64))
    assert struct.unpack_from("<I", img, 4)[0] == len(sizes)


@pytest.mark.gpu
def test_gpu_every_last_chunk_count(pvvc, oracle, tmp_path):
    img = _frame(pvvc, np.random.default_rng(4))
    path = tmp_path / "residues_0.frame"
    path.write_bytes(img)
    ref = oracle.StreamDecoder(fp64=False).decode(oracle.OracleFrame(str(path)), 4)
    assert (ref["status"] == 0).all()
    f = pvvc.Frame.load(str(path))
    S = f.slice_cnt
    mx = int(f.meta()["size"].max())
    for waves in (2, 1, 4):
        _say("residues drop-in, waves", waves)
        assert pvvc.hip_lib().pvvc_set_dropin_tuning(waves, 32 * 1024) == 0
        pts = pvvc.decode_dropin(f, pvvc.State(S, mx))
        np.testing.assert_array_equal(np.stack([pts["x"], pts["y"], pts["z"]], 1), ref["xyz"])
        np.testing.assert_array_equal(np.stack([pts["r"], pts["g"], pts["b"]], 1), ref["rgbf"] / np.float32(255))
    assert pvvc.hip_lib().pvvc_set_dropin_tuning(4, 32 * 1024) == 0  # the default
    for mode in (pvvc.MODE_PER_FRAME, pvvc.MODE_PERSISTENT, pvvc.MODE_GRAPH):
        _say("residues native mode", mode)
        dec = pvvc.Decoder(S, mx)
        b = dec.batch([f])
        xyz, rgb = dec.decode_host(b, mode, pvvc.OUT_NATIVE)
        assert (b.status() == 0).all()
        np.testing.assert_array_equal(xyz, ref["xyz"])
        np.testing.assert_array_equal(rgb, ref["rgb"])
