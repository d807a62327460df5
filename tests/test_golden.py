"""Golden fixtures (tests/golden/, written by make_golden.py): a hand-built
intra + predictive stream whose expected output comes from FP64 closed forms
(two-leaf Haar pair, single leaf, integer motion, BT.601).  No channel lies
within 0.05 of a rounding boundary, so FP32 decoders must match exactly."""
import json
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def golden():
    d = json.loads((GOLDEN / "golden_expected.json").read_text())
    paths = [GOLDEN / f for f in d["frames"]]
    exp = [(np.array(e["xyz"], np.float32), np.array(e["rgb"], np.uint8)) for e in d["expected"]]
    return paths, exp


def test_golden_oracle(oracle, golden):
    paths, exp = golden
    dec = oracle.StreamDecoder(fp64=True)
    for p, (xyz, rgb) in zip(paths, exp):
        out = dec.decode(oracle.OracleFrame(str(p)), 2)
        assert (out["status"] == 0).all()
        np.testing.assert_array_equal(out["xyz"], xyz)
        np.testing.assert_array_equal(out["rgb"], rgb)
        # FP64 path: within 0.05 of the closed form before rounding
        assert np.abs(np.round(out["rgb64"]) - rgb).max() == 0


def test_golden_container(pvvc, golden):
    paths, exp = golden
    f0 = pvvc.Frame.load(str(paths[0]))
    m = f0.meta()
    assert m["index"].tolist() == [0, 1, 2] and m["size"].tolist() == [2, 2, 1]
    assert m["geometry_size"].tolist() == [26, 26, 25]   # patch 2 unwrapped from zstd
    assert f0.points == len(exp[0][0])


@pytest.mark.gpu
def test_golden_gpu_native(pvvc, golden):
    paths, exp = golden
    frames = [pvvc.Frame.load(str(p)) for p in paths]
    for mode in (pvvc.MODE_PER_FRAME, pvvc.MODE_PERSISTENT, pvvc.MODE_GRAPH):
        dec = pvvc.Decoder(3, 2)
        b = dec.batch(frames)
        xyz, rgb = dec.decode_host(b, mode, pvvc.OUT_NATIVE)
        assert (b.status() == 0).all()
        for k, (exyz, ergb) in enumerate(exp):
            o, n = b.frame_offset[k], b.frame_points[k]
            np.testing.assert_array_equal(xyz[o:o + n], exyz)
            np.testing.assert_array_equal(rgb[o:o + n], ergb)


@pytest.mark.gpu
def test_golden_gpu_dropin(pvvc, golden):
    paths, exp = golden
    st = pvvc.State(3, 2)
    for p, (exyz, ergb) in zip(paths, exp):
        pts = pvvc.decode_dropin(pvvc.Frame.load(str(p)), st)
        xyz = np.stack([pts["x"], pts["y"], pts["z"]], 1)
        np.testing.assert_array_equal(xyz, exyz)
        rgb = np.round(np.stack([pts["r"], pts["g"], pts["b"]], 1).astype(np.float64) * 255.0)
        np.testing.assert_array_equal(rgb, ergb.astype(np.float64))
