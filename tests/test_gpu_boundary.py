"""GPU tests of the C-ABI boundary's object model and error reporting.

* one Decoder reused across persistent batches whose per-item workspace
  shrinks while the grid grows (the slab stride is this call's bound);
* the captured hipGraph is re-made when its key changes (launch shape,
  white-small setting);
* launch_cudaProcess reports an out-of-range index[s] and failed slices
  through pvvc_state_synchronize / pvvc_last_error;
* two host threads, each with its own state, decoder and stream, decode two
  different streams concurrently, both bit-exact against the oracle;
* a predictive slice that carries a geometry blob decodes on its patch's
  intra tree, as the reference does (invert_raht_octree.cc:71-78);
* a GoP read from .slice files decodes like its .frame files.
"""
import ctypes as C
import sys
import threading

import numpy as np
import pytest
from _frames import pack_frame, unpack_frame

pytestmark = pytest.mark.gpu

INVALID_ARG, EMPTY_REFERENCE = 8, 2


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dims(frames):
    metas = [f.meta() for f in frames]
    return max(int(m["index"].max()) for m in metas) + 1, max(int(m["size"].max()) for m in metas)


def _oracle_refs(oracle, tmp_path, name, imgs, fp64=False):
    od = oracle.StreamDecoder(fp64=fp64)
    refs = []
    for k, b in enumerate(imgs):
        p = tmp_path / f"{name}_{k}.frame"
        p.write_bytes(b)
        refs.append(od.decode(oracle.OracleFrame(str(p)), 8))
    return refs


def _ws_bound(frames):
    """Largest per-item workspace bound of a batch (item_ws_bound of the
    library, 256-B aligned), intra items only (these streams have no P-frames)."""
    need = 0
    for f in frames:
        m = f.meta()
        for n, g in zip(m["size"].astype(np.int64), m["geometry_size"].astype(np.int64)):
            need = max(need, 12 * n + 2 * (4 * (g + 1) + 16) + 48 * n + 512)
    return (need + 255) & ~255


def test_decoder_reuse_shrinking_items_persistent(pvvc, oracle, tmp_path):
    big = pvvc.synth_stream("c0", n_frames=1, seed=7, threads=8, patch_min=40000, patch_max=40000)
    small = pvvc.synth_stream("c0", n_frames=1, seed=7, threads=8, patch_min=100, patch_max=200)
    fb, fs = [pvvc.Frame.parse(b) for b in big], [pvvc.Frame.parse(b) for b in small]
    cu = pvvc.device_info()["cu"]
    # the first batch's slab (auto shape for small segments: grid = items, up to 4 per CU)
    grid_b, grid_s = min(fb[0].slice_cnt, 4 * cu), min(fs[0].slice_cnt, 4 * cu)
    slab_b = _ws_bound(fb) * grid_b
    _say(f"big: {fb[0].slice_cnt} items, bound {_ws_bound(fb)} B, slab {slab_b} B; small: {fs[0].slice_cnt} items, "
         f"bound {_ws_bound(fs)} B, grid {grid_s}")
    # the case under test: the second batch fits the first slab with its own
    # stride, and would not with the first batch's stride
    assert _ws_bound(fs) * grid_s <= slab_b < _ws_bound(fb) * grid_s
    np_, mp = _dims(fb + fs)
    dec = pvvc.Decoder(np_, mp)
    for name, imgs, frames in (("big", big, fb), ("small", small, fs), ("big2", big, fb)):
        _say("persistent decode of", name)
        refs = _oracle_refs(oracle, tmp_path, name, imgs)
        b = dec.batch(frames)
        xyz, rgb = dec.decode_host(b, pvvc.MODE_PERSISTENT, pvvc.OUT_NATIVE)
        assert (b.status() == 0).all()
        np.testing.assert_array_equal(xyz, refs[0]["xyz"])
        np.testing.assert_array_equal(rgb, refs[0]["rgb"])


def test_graph_key_follows_shape_and_white_small(pvvc, oracle, tmp_path):
    imgs = pvvc.synth_stream("c0", n_frames=2, seed=9, threads=8, patch_min=4, patch_max=16)
    frames = [pvvc.Frame.parse(b) for b in imgs]
    refs = _oracle_refs(oracle, tmp_path, "tiny", imgs)
    sizes = np.concatenate([np.repeat(f.meta()["size"], f.meta()["size"]) for f in frames])
    small = sizes < 10
    _say(f"tiny patches: {frames[0].slice_cnt} per frame, {int(small.sum())} points in patches under 10")
    assert small.any() and (~small).any()
    rgbf = np.concatenate([r["rgbf"] for r in refs]) / np.float32(255)
    dec = pvvc.Decoder(*_dims(frames))
    b = dec.batch(frames)

    def run():
        pts = dec.decode_host(b, pvvc.MODE_GRAPH, pvvc.OUT_POINTS)
        assert (b.status() == 0).all()
        return np.stack([pts["r"], pts["g"], pts["b"]], 1)

    np.testing.assert_array_equal(run(), rgbf)
    dec.set_tuning(waves_per_patch=4, lds_bytes=16384)
    np.testing.assert_array_equal(run(), rgbf)
    dec.set_white_small_patches(1)
    col = run()
    assert (col[small] == 1.0).all()  # the re-captured graph applies the new setting
    np.testing.assert_array_equal(col[~small], rgbf[~small])
    dec.set_white_small_patches(0)
    np.testing.assert_array_equal(run(), rgbf)


def test_dropin_reports_bad_index_and_failed_slices(pvvc, oracle, c0_stream, tmp_path):
    frames = [pvvc.Frame.parse(b) for b in c0_stream[:2]]
    refs = _oracle_refs(oracle, tmp_path, "c0", c0_stream[:2])
    n_patches, max_pts = _dims(frames)
    st = pvvc.State(n_patches, max_pts)
    assert st.synchronize() == 0
    staged = pvvc.StagedFrame(frames[0])
    bad = np.array([n_patches + 5], np.int32)
    _check = pvvc.hip_lib().pvvc_memcpy_h2d(C.c_void_p(staged.d.index), bad.ctypes.data, 4)
    assert _check == 0
    out = pvvc.DeviceBuffer(24 * frames[0].points)
    pvvc.hip_lib().pvvc_last_error()  # clear this thread's record
    pvvc.launch_dropin(staged, st, out)
    assert st.synchronize() == INVALID_ARG
    assert pvvc.hip_lib().pvvc_last_error() == INVALID_ARG
    assert st.synchronize() == 0  # read and cleared
    pts = out.read(pvvc.POINTS_DTYPE, frames[0].points)
    n0 = int(frames[0].meta()["size"][0])
    xyz = np.stack([pts["x"], pts["y"], pts["z"]], 1)
    np.testing.assert_array_equal(xyz[n0:], refs[0]["xyz"][n0:])  # every other slice decoded
    # a predictive frame on a fresh state: no reference
    st2 = pvvc.State(n_patches, max_pts)
    pvvc.launch_dropin(pvvc.StagedFrame(frames[1]), st2, out)
    assert st2.synchronize() == EMPTY_REFERENCE
    assert pvvc.hip_lib().pvvc_last_error() == EMPTY_REFERENCE


def test_two_threads_own_state_decoder_and_stream(pvvc, oracle, tmp_path):
    streams = {}
    for seed in (7, 13):
        imgs = pvvc.synth_stream("c0", n_frames=8, seed=seed, threads=8)
        streams[seed] = (imgs, _oracle_refs(oracle, tmp_path, f"s{seed}", imgs))
    barrier = threading.Barrier(2)
    errors = {}

    def worker(seed):
        try:
            imgs, refs = streams[seed]
            frames = [pvvc.Frame.parse(b) for b in imgs]
            s = pvvc.Stream()
            st = pvvc.State(*_dims(frames))
            st.set_stream(s)
            dec = pvvc.Decoder(*_dims(frames))
            dec.set_stream(s)
            out = pvvc.DeviceBuffer(24 * max(f.points for f in frames))
            barrier.wait()
            for rep in range(3):
                for k, (f, ref) in enumerate(zip(frames, refs)):
                    staged = pvvc.StagedFrame(f)
                    pvvc.launch_dropin(staged, st, out)
                    assert st.synchronize() == 0
                    pts = out.read(pvvc.POINTS_DTYPE, f.points)
                    np.testing.assert_array_equal(np.stack([pts["x"], pts["y"], pts["z"]], 1), ref["xyz"])
                    np.testing.assert_array_equal(np.stack([pts["r"], pts["g"], pts["b"]], 1),
                                                  ref["rgbf"] / np.float32(255))
                b = dec.batch(frames)
                o = dec.outputs(b, pvvc.OUT_NATIVE)
                dec.decode(b, pvvc.MODE_PERSISTENT, o)
                s.synchronize()
                assert (b.status() == 0).all()
                n = b.total_points
                rgb = o.rgb.read(np.uint8, 3 * n).reshape(n, 3)
                for k, ref in enumerate(refs):
                    np.testing.assert_array_equal(rgb[b.frame_offset[k]:b.frame_offset[k] + b.frame_points[k]],
                                                  ref["rgb"])
            assert pvvc.hip_lib().pvvc_last_error() == 0
            # everything that synchronises on the stream goes before it
            b.free()
            dec.close()
            st.free()
            s.close()
        except BaseException as e:  # reported by the main thread
            errors[seed] = e

    th = [threading.Thread(target=worker, args=(s,)) for s in streams]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=90)
    assert not any(t.is_alive() for t in th)
    assert not errors, errors


def test_predictive_slice_geometry_blob_ignored(pvvc, oracle, c0_stream, tmp_path):
    ts0, s0 = unpack_frame(pvvc, c0_stream[0])
    ts1, s1 = unpack_frame(pvvc, c0_stream[1])
    # predictive slices carrying foreign geometry: another patch's tree, and a
    # blob shorter than a tree header
    s1[0]["geo"] = s0[5]["geo"]
    s1[1]["geo"] = b"\x01\x02\x03"
    imgs = [pack_frame(ts0, s0), pack_frame(ts1, s1)]
    refs = _oracle_refs(oracle, tmp_path, "pgeo", imgs)
    plain = _oracle_refs(oracle, tmp_path, "plain", c0_stream[:2])
    for r, q in zip(refs, plain):
        assert (r["status"] == 0).all()
        np.testing.assert_array_equal(r["rgb"], q["rgb"])
    frames = [pvvc.Frame.parse(b) for b in imgs]
    st = pvvc.State(*_dims(frames))
    for f, ref in zip(frames, refs):
        pts = pvvc.decode_dropin(f, st)
        np.testing.assert_array_equal(np.stack([pts["x"], pts["y"], pts["z"]], 1), ref["xyz"])
        np.testing.assert_array_equal(np.stack([pts["r"], pts["g"], pts["b"]], 1), ref["rgbf"] / np.float32(255))
    for mode in (pvvc.MODE_PER_FRAME, pvvc.MODE_PERSISTENT):
        dec = pvvc.Decoder(*_dims(frames))
        b = dec.batch(frames)
        xyz, rgb = dec.decode_host(b, mode, pvvc.OUT_NATIVE)
        assert (b.status() == 0).all()
        for k, ref in enumerate(refs):
            o, n = b.frame_offset[k], b.frame_points[k]
            np.testing.assert_array_equal(rgb[o:o + n], ref["rgb"])


def test_gop_from_slice_files_matches_oracle(pvvc, oracle, c0_stream, tmp_path):
    """A c0 GoP written as .slice files (one per patch and frame), read back
    with pvvc_load_slices (LoadSlice + ChangeSliceToFrame, the reference
    decoder harness's input, src/codec/pvvc_decompression.cc:34-84), decodes
    in one persistent launch bit-exactly like the oracle decodes the .frame
    files."""
    from test_slice_io import pack_slice, split_frame
    refs = _oracle_refs(oracle, tmp_path, "c0s", c0_stream)
    frames = []
    for k, image in enumerate(c0_stream):
        ts, slices = split_frame(image)
        paths = []
        for s in slices:
            p = tmp_path / f"f{k}_time_{ts:04d}_slice_{s['index']:04d}.slice"
            p.write_bytes(pack_slice(s["type"], ts, s["index"], s["size"], s["qp"], s["mv"], s["geo"] or None,
                                     s["col"]))
            paths.append(p)
        frames.append(pvvc.Frame.from_slices(paths[::-1]))
    _say("GoP from", sum(f.slice_cnt for f in frames), "slice files")
    dec = pvvc.Decoder(*_dims(frames))
    b = dec.batch(frames)
    xyz, rgb = dec.decode_host(b, pvvc.MODE_PERSISTENT, pvvc.OUT_NATIVE)
    assert (b.status() == 0).all()
    for k, ref in enumerate(refs):
        o, n = b.frame_offset[k], b.frame_points[k]
        np.testing.assert_array_equal(xyz[o:o + n], ref["xyz"])
        np.testing.assert_array_equal(rgb[o:o + n], ref["rgb"])
