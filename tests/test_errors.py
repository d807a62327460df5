"""Error paths against the oracle: single-fault slices of an intra and a
predictive frame report the oracle's status in every decode mode, and the
healthy slices around them still decode bit-exactly."""
import sys

import numpy as np
import pytest
from _frames import bits_to_bytes, pack_frame, unpack_frame

OK, BAD_SLICE, EMPTY_REFERENCE, UNMATCHED, BITSTREAM, SYMBOL_RANGE = 0, 1, 2, 3, 5, 6


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _symbol_range_blob():
    """k = 0 symbols of 31 one-bits raise k_RP 8 -> 38 -> 68 -> 98 -> 128 (k_R
    2, 9, 17, 24, 32); the fifth is 31 << 32, outside int32."""
    bits = ""
    for kr in (2, 9, 17, 24, 32):
        bits += "1" * 31 + "0" + "0" * kr
    return bits_to_bytes(bits + "0" * 64)


def _bad_run_blob():
    """Two zero symbols take k_P to 6 (run mode); then only one-bits: the run
    exponent passes 30."""
    return bits_to_bytes("00000" + "1" * 200)


@pytest.fixture(scope="module")
def faulty(pvvc, c0_stream, tmp_path_factory):
    ts0, s0 = unpack_frame(pvvc, c0_stream[0])
    ts1, s1 = unpack_frame(pvvc, c0_stream[1])
    assert all(s["type"] == 1 for s in s0) and all(s["type"] == 3 for s in s1)
    by_idx1 = {s["index"]: i for i, s in enumerate(s1)}
    expect0, expect1 = {}, {}

    def mut0(i, status, **kw):
        s0[i].update(kw)
        expect0[i] = status

    g = lambda i: bytearray(s0[i]["geo"])  # noqa: E731
    mut0(0, BITSTREAM, col=s0[0]["col"][: len(s0[0]["col"]) // 3])
    mut0(1, UNMATCHED, size=s0[1]["size"] + 1)
    mut0(2, BAD_SLICE, type=0x00)
    geo3 = g(3)
    geo3[24] = 0
    mut0(3, BAD_SLICE, geo=bytes(geo3))
    mut0(4, BAD_SLICE, geo=s0[4]["geo"][:30])
    geo5 = g(5)
    geo5[25] = 0  # root occupancy byte
    mut0(5, BAD_SLICE, geo=bytes(geo5))
    mut0(6, BITSTREAM, col=_bad_run_blob())
    mut0(7, SYMBOL_RANGE, col=_symbol_range_blob())
    # predictive frame: the patches whose intra slice failed have no reference
    # (slice 2, an invalid type, leaves its never-filled state untouched)
    for i in range(8):
        j = by_idx1.get(s0[i]["index"])
        if j is not None:
            expect1[j] = EMPTY_REFERENCE
    healthy = [j for j in range(len(s1)) if j not in expect1]
    j_size, j_col = healthy[0], healthy[1]
    s1[j_size]["size"] += 1
    expect1[j_size] = UNMATCHED
    s1[j_col]["col"] = s1[j_col]["col"][: len(s1[j_col]["col"]) // 4]
    expect1[j_col] = BITSTREAM
    d = tmp_path_factory.mktemp("faulty")
    paths = []
    for k, (ts, sl) in enumerate(((ts0, s0), (ts1, s1))):
        p = d / f"faulty_{k}.frame"
        p.write_bytes(pack_frame(ts, sl))
        paths.append(p)
    return paths, (expect0, expect1)


def _oracle(oracle, paths):
    dec = oracle.StreamDecoder(fp64=False)
    return [dec.decode(oracle.OracleFrame(str(p)), 8) for p in paths]


def _check_frame(k, status, ref, expect, xyz=None, rgb=None, offsets=None):
    want = np.zeros(len(ref["status"]), np.int32)
    for i, v in expect.items():
        want[i] = v
    np.testing.assert_array_equal(ref["status"], want, err_msg=f"oracle statuses, frame {k}")
    np.testing.assert_array_equal(np.asarray(status), want, err_msg=f"GPU statuses, frame {k}")
    if xyz is None:
        return
    sizes = ref["meta"]["size"].astype(np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)])
    for s in np.flatnonzero(want == OK):
        a, b = off[s], off[s + 1]
        np.testing.assert_array_equal(xyz[offsets + a:offsets + b], ref["xyz"][a:b])
        np.testing.assert_array_equal(rgb[offsets + a:offsets + b], ref["rgb"][a:b])


def test_oracle_statuses_match_expectation(oracle, faulty):
    paths, expect = faulty
    for k, ref in enumerate(_oracle(oracle, paths)):
        _check_frame(k, ref["status"], ref, expect[k])


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1, 2], ids=["per_frame", "persistent", "graph"])
def test_native_statuses_and_healthy_slices(pvvc, oracle, faulty, mode):
    paths, expect = faulty
    refs = _oracle(oracle, paths)
    frames = [pvvc.Frame.load(str(p)) for p in paths]
    n_patches = max(int(f.meta()["index"].max()) for f in frames) + 1
    max_pts = max(int(f.meta()["size"].max()) for f in frames) + 1
    _say("faulty native mode", mode)
    dec = pvvc.Decoder(n_patches, max_pts)
    b = dec.batch(frames)
    xyz, rgb = dec.decode_host(b, mode, pvvc.OUT_NATIVE)
    st = b.status()
    first = 0
    for k, f in enumerate(frames):
        _check_frame(k, st[first:first + f.slice_cnt], refs[k], expect[k], xyz, rgb, b.frame_offset[k])
        first += f.slice_cnt


@pytest.mark.gpu
def test_native_statuses_one_frame_per_batch(pvvc, oracle, faulty):
    """Intra and predictive frames in separate persistent launches: the
    predictive slices read the resident state of the failed intra slices."""
    paths, expect = faulty
    refs = _oracle(oracle, paths)
    frames = [pvvc.Frame.load(str(p)) for p in paths]
    n_patches = max(int(f.meta()["index"].max()) for f in frames) + 1
    max_pts = max(int(f.meta()["size"].max()) for f in frames) + 1
    dec = pvvc.Decoder(n_patches, max_pts)
    for k, f in enumerate(frames):
        b = dec.batch([f])
        xyz, rgb = dec.decode_host(b, pvvc.MODE_PERSISTENT, pvvc.OUT_NATIVE)
        _check_frame(k, b.status(), refs[k], expect[k], xyz, rgb, 0)


@pytest.mark.gpu
def test_dropin_statuses_and_healthy_slices(pvvc, oracle, faulty):
    paths, expect = faulty
    refs = _oracle(oracle, paths)
    frames = [pvvc.Frame.load(str(p)) for p in paths]
    n_patches = max(int(f.meta()["index"].max()) for f in frames) + 1
    max_pts = max(int(f.meta()["size"].max()) for f in frames) + 1
    state = pvvc.State(n_patches, max_pts)
    for k, f in enumerate(frames):
        _say("faulty drop-in frame", k)
        pts = pvvc.decode_dropin(f, state)
        idx = f.meta()["index"]
        st = [state.status(int(i)) for i in idx]
        xyz = np.stack([pts["x"], pts["y"], pts["z"]], 1)
        rgbf = np.stack([pts["r"], pts["g"], pts["b"]], 1)
        _check_frame(k, st, refs[k], expect[k])
        sizes = refs[k]["meta"]["size"].astype(np.int64)
        off = np.concatenate([[0], np.cumsum(sizes)])
        want = np.array([expect[k].get(i, OK) for i in range(f.slice_cnt)])
        for s in np.flatnonzero(want == OK):
            a, c = off[s], off[s + 1]
            np.testing.assert_array_equal(xyz[a:c], refs[k]["xyz"][a:c])
            np.testing.assert_array_equal(rgbf[a:c], refs[k]["rgbf"][a:c] / np.float32(255))


@pytest.mark.gpu
def test_one_call_decodes_report_statuses(pvvc, oracle, faulty):
    """pvvc_decode_frame / pvvc_decode_gop: per-slice statuses equal the
    oracle's, healthy slices decode bit-exactly, and the return code is the
    first failing slice's status."""
    paths, expect = faulty
    refs = _oracle(oracle, paths)
    frames = [pvvc.Frame.load(str(p)) for p in paths]
    n_patches = max(int(f.meta()["index"].max()) for f in frames) + 1
    max_pts = max(int(f.meta()["size"].max()) for f in frames) + 1
    _say("one-call decode_frame")
    dec = pvvc.Decoder(n_patches, max_pts)
    for k, f in enumerate(frames):
        out = dec.outputs_for([f])
        rc, st = dec.decode_frame(f, out)
        n = f.points
        xyz = out.xyz.read(np.float32, 3 * n).reshape(n, 3)
        rgb = out.rgb.read(np.uint8, 3 * n).reshape(n, 3)
        _check_frame(k, st, refs[k], expect[k], xyz, rgb, 0)
        assert rc == st[np.flatnonzero(st)[0]]
    _say("one-call decode_gop")
    dec = pvvc.Decoder(n_patches, max_pts)
    out = dec.outputs_for(frames)
    rc, st = dec.decode_gop(frames, out)
    n = sum(f.points for f in frames)
    xyz = out.xyz.read(np.float32, 3 * n).reshape(n, 3)
    rgb = out.rgb.read(np.uint8, 3 * n).reshape(n, 3)
    first, off = 0, 0
    for k, f in enumerate(frames):
        _check_frame(k, st[first:first + f.slice_cnt], refs[k], expect[k], xyz, rgb, off)
        first += f.slice_cnt
        off += f.points
    assert rc == st[np.flatnonzero(st)[0]]


@pytest.mark.gpu
def test_one_call_gop_equals_batch_api(pvvc, c0_stream):
    frames = [pvvc.Frame.parse(b) for b in c0_stream]
    n_patches = max(int(f.meta()["index"].max()) for f in frames) + 1
    max_pts = max(int(f.meta()["size"].max()) for f in frames) + 1
    _say("one-call gop vs batch, c0")
    dec = pvvc.Decoder(n_patches, max_pts)
    xyz_b, rgb_b = dec.decode_host(dec.batch(frames), pvvc.MODE_PERSISTENT, pvvc.OUT_NATIVE)
    dec2 = pvvc.Decoder(n_patches, max_pts)
    out = dec2.outputs_for(frames, pvvc.OUT_POINTS)
    rc, st = dec2.decode_gop(frames, out)
    assert rc == 0 and (st == 0).all()
    pts = out.pts.read(pvvc.POINTS_DTYPE, len(xyz_b))
    np.testing.assert_array_equal(np.stack([pts["x"], pts["y"], pts["z"]], 1), xyz_b)
    # Points carries unrounded, unclamped float/255 colours; the native output
    # rounds and saturates them
    rgb_p = np.clip(np.stack([pts["r"], pts["g"], pts["b"]], 1) * 255, 0, 255)
    assert np.abs(rgb_p - rgb_b.astype(np.float32)).max() <= 0.5 + 1e-3
    dec3 = pvvc.Decoder(n_patches, max_pts)
    for k, f in enumerate(frames):
        out = dec3.outputs_for([f])
        rc, st = dec3.decode_frame(f, out)
        assert rc == 0 and (st == 0).all()
        o = int(sum(g.points for g in frames[:k]))
        np.testing.assert_array_equal(out.xyz.read(np.float32, 3 * f.points).reshape(-1, 3), xyz_b[o:o + f.points])
        np.testing.assert_array_equal(out.rgb.read(np.uint8, 3 * f.points).reshape(-1, 3), rgb_b[o:o + f.points])
