"""GPU tests of the serving loop: continuous batching over many streams,
HBM-footprint admission, resident GoP state, and that a stream's decoded
frames are the same alone and in the mix (bit-exact: the decode of a slice
does not depend on what else shares its launch)."""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def streams(pvvc):
    out = {}
    for seed in (7, 11, 13):
        imgs = pvvc.synth_stream("c0", n_frames=8, seed=seed, threads=8)
        out[seed] = ([pvvc.Frame.parse(b) for b in imgs], imgs)
    return out


def _dims(frames):
    metas = [f.meta() for f in frames]
    return (max(int(m["index"].max()) for m in metas) + 1, max(int(m["size"].max()) for m in metas),
            max(f.points for f in frames))


def _drain(srv, want, timeout=60.0, keep=None):
    """Polls `want` completions.  Output slots are handed back as they come
    (each stream has a small ring, so holding them would stall it); with
    `keep`, each completion's output is read first into keep[(stream, seq)]."""
    got = []
    t0 = time.time()
    while len(got) < want:
        for c in srv.poll(256, 100):
            if keep is not None:
                keep[(c["stream_id"], c["seq"])] = srv.read(c)
            srv.release(c)
            got.append(c)
        assert time.time() - t0 < timeout, f"only {len(got)} of {want} completions"
    return got


def _decode_alone(pvvc, frames):
    srv = pvvc.Server(lanes=2)
    sid = srv.open_stream(*_dims(frames))
    assert srv.stream_state(sid) == pvvc.STREAM_ACTIVE
    res = []
    for k, f in enumerate(frames):
        srv.submit(sid, f)
        keep = {}
        [c] = _drain(srv, 1, keep=keep)
        assert c["status"] == 0, c
        res.append(keep[(sid, k)])
    srv.close()
    return res


def test_serving_alone_matches_oracle_and_mix(pvvc, oracle, streams, tmp_path):
    frames_a, imgs_a = streams[7]
    _say("serving alone: stream seed 7")
    alone = _decode_alone(pvvc, frames_a)

    dec = oracle.StreamDecoder(fp64=False)
    for k, img in enumerate(imgs_a):
        p = tmp_path / f"a_{k}.frame"
        p.write_bytes(img)
        ref = dec.decode(oracle.OracleFrame(str(p)), 8)
        np.testing.assert_array_equal(alone[k][0], ref["xyz"])
        np.testing.assert_array_equal(alone[k][1], ref["rgb"])

    _say("serving mix: 5 streams interleaved")
    srv = pvvc.Server(lanes=3)
    plan = [(7, streams[7][0]), (11, streams[11][0]), (13, streams[13][0]), (11, streams[11][0]),
            (13, streams[13][0])]
    ids = [srv.open_stream(*_dims(fr)) for _, fr in plan]
    for k in range(8):
        for sid, (_, fr) in zip(ids, plan):
            srv.submit(sid, fr[k], tag=k)
    by = {}
    got = _drain(srv, 8 * len(plan), keep=by)
    assert all(c["status"] == 0 for c in got), [c for c in got if c["status"]][:3]
    st = srv.stats()
    _say("  stats", st)
    assert st["frames_done"] == 8 * len(plan)
    assert st["batches"] < st["frames_done"]  # frames of several streams shared launches
    for k in range(8):
        np.testing.assert_array_equal(by[(ids[0], k)][0], alone[k][0])
        np.testing.assert_array_equal(by[(ids[0], k)][1], alone[k][1])
    # the two copies of each other stream decode identically as well
    for k in range(8):
        for a, b in ((1, 3), (2, 4)):
            np.testing.assert_array_equal(by[(ids[a], k)][1], by[(ids[b], k)][1])
    srv.close()


def test_serving_admission_by_footprint(pvvc, streams):
    fa, _ = streams[7]
    fb, _ = streams[11]
    probe = pvvc.Server(lanes=1)
    a = probe.open_stream(*_dims(fa))
    used_a = probe.stats()["hbm_used"]
    b = probe.open_stream(*_dims(fb), out_ring=8)
    used_ab = probe.stats()["hbm_used"]
    probe.close()
    assert used_a > 0 and used_ab > used_a

    _say("admission: budget", used_ab - 1)
    srv = pvvc.Server(hbm_budget=used_ab - 1, lanes=2)
    a = srv.open_stream(*_dims(fa))
    b = srv.open_stream(*_dims(fb), out_ring=8)
    assert srv.stream_state(a) == pvvc.STREAM_ACTIVE
    assert srv.stream_state(b) == pvvc.STREAM_WAITING
    assert srv.stats()["streams_waiting"] == 1
    for f in fb:  # queued while the stream waits for memory
        srv.submit(b, f)
    for f in fa:
        srv.submit(a, f)
    got_a = _drain(srv, 8)
    assert all(c["stream_id"] == a and c["status"] == 0 for c in got_a)
    time.sleep(0.05)
    assert not srv.poll(16, 0)  # nothing of b decoded before admission
    srv.close_stream(a)
    assert srv.stream_state(b) == pvvc.STREAM_ACTIVE
    keep = {}
    got_b = _drain(srv, 8, keep=keep)
    assert all(c["stream_id"] == b and c["status"] == 0 for c in got_b), got_b
    out_b = [keep[(b, k)] for k in range(8)]
    srv.close()
    # all 8 frames of b waited, so intra and predictive slices shared launches;
    # compare with the batched decoder on the same frames
    n_p, max_pts, _ = _dims(fb)
    dec = pvvc.Decoder(n_p, max_pts)
    bt = dec.batch(fb)
    xyz, rgb = dec.decode_host(bt, pvvc.MODE_PERSISTENT, pvvc.OUT_NATIVE)
    for k in range(8):
        o, n = bt.frame_offset[k], bt.frame_points[k]
        np.testing.assert_array_equal(out_b[k][0], xyz[o:o + n])
        np.testing.assert_array_equal(out_b[k][1], rgb[o:o + n])


def test_serving_predictive_before_intra_reports(pvvc, streams):
    fa, _ = streams[7]
    srv = pvvc.Server(lanes=1)
    sid = srv.open_stream(*_dims(fa))
    srv.submit(sid, fa[1])
    [c] = _drain(srv, 1)
    assert c["status"] == 2  # EMPTY_REFERENCE: no intra slice yet
    srv.submit(sid, fa[0])
    srv.submit(sid, fa[1])
    got = sorted(_drain(srv, 2), key=lambda c: c["seq"])
    assert [c["status"] for c in got] == [0, 0]
    srv.close()


def test_serving_poisson_load_outputs(pvvc, streams):
    """Open-loop Poisson load over 8 streams: every frame completes and each
    output equals the stream's alone decode of the same GoP position."""
    from pvvc import serve
    fa, _ = streams[7]
    fb, _ = streams[11]
    ref = {7: _decode_alone(pvvc, fa), 11: _decode_alone(pvvc, fb)}
    specs = [serve.StreamSpec("a", fa), serve.StreamSpec("b", fb)] * 4
    keys = [7, 11] * 4
    bad = []
    srv = pvvc.Server(lanes=3)

    def check(i, k, c):
        xyz, rgb = srv.read(c)
        e = ref[keys[i]][k % 8]
        if c["status"] != 0 or not (np.array_equal(xyz, e[0]) and np.array_equal(rgb, e[1])):
            bad.append((i, k, c["status"]))

    _say("poisson load: 8 streams, 60 fps, 1 s")
    st = serve.run(srv, specs, fps=60.0, duration=1.0, seed=3, check=check)
    srv.close()
    _say("  stats", st)
    assert st["frames_completed"] == st["frames_offered"] > 0
    assert st["frames_failed"] == 0
    assert not bad, bad[:5]
