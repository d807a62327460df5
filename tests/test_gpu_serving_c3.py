"""Serving at the mixed-density shape: 64 concurrent streams, each with its
own seed -- 21 at ~100k points / 9-bit, 21 at ~400k / 10-bit, 22 at ~1M /
10-bit, all GoP 8, qp 4 -- under Poisson arrivals at 30 fps per stream.

Every completion must equal that stream's decode alone (the same GoP on a
decoder of its own, one persistent launch per GoP), bit for bit; one stream
of each density is checked against the CPU oracle as well.  Outputs are
compared by 64-bit hashes of their xyz and RGB bytes, so that 64 streams'
decodes need not be held in host memory.
"""
import sys
import time

import numpy as np
import pytest
import xxhash

pytestmark = pytest.mark.gpu

KINDS = [("c3_100k", 21), ("c3_400k", 21), ("c3_1m", 22)]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _digest(xyz: np.ndarray, rgb: np.ndarray) -> tuple[int, int]:
    return xxhash.xxh3_64_intdigest(np.ascontiguousarray(xyz).data), xxhash.xxh3_64_intdigest(
        np.ascontiguousarray(rgb).data)


def _dims(frames):
    metas = [f.meta() for f in frames]
    return max(int(m["index"].max()) for m in metas) + 1, max(int(m["size"].max()) for m in metas)


@pytest.mark.timeout(900)
def test_serving_64_mixed_streams_equal_alone_and_oracle(pvvc, oracle, tmp_path):
    from pvvc import serve
    t0 = time.time()
    specs, alone = [], []
    seed = 1000
    for kind, count in KINDS:
        gops = []
        for i in range(count):
            imgs = pvvc.synth_stream(kind, n_frames=8, seed=seed, threads=16)
            gops.append((seed, imgs, [pvvc.Frame.parse(b) for b in imgs]))
            seed += 1
        _say(f"{kind}: {count} streams generated ({time.time() - t0:.1f} s), "
             f"{gops[0][2][0].points} points in the first frame of the first")
        np_ = max(_dims(g[2])[0] for g in gops)
        mp = max(_dims(g[2])[1] for g in gops)
        dec = pvvc.Decoder(np_, mp)
        for j, (sd, imgs, frames) in enumerate(gops):
            b = dec.batch(frames)
            xyz, rgb = dec.decode_host(b, pvvc.MODE_PERSISTENT, pvvc.OUT_NATIVE)
            assert (b.status() == 0).all(), (kind, sd)
            sl = [slice(b.frame_offset[k], b.frame_offset[k] + b.frame_points[k]) for k in range(8)]
            alone.append([_digest(xyz[s], rgb[s]) for s in sl])
            if j == 0:  # one stream per density against the oracle
                od = oracle.StreamDecoder(fp64=False)
                for k, img in enumerate(imgs):
                    p = tmp_path / f"{kind}_{k}.frame"
                    p.write_bytes(img)
                    ref = od.decode(oracle.OracleFrame(str(p)), 16)
                    assert (ref["status"] == 0).all()
                    np.testing.assert_array_equal(xyz[sl[k]], ref["xyz"], err_msg=f"{kind} frame {k}")
                    np.testing.assert_array_equal(rgb[sl[k]], ref["rgb"], err_msg=f"{kind} frame {k}")
                _say(f"  {kind} seed {sd}: 8 frames equal to the oracle")
            b.free()
            specs.append(serve.StreamSpec(kind, frames))
        dec.close()
    _say(f"alone decodes done ({time.time() - t0:.1f} s); serving 64 streams")
    srv = pvvc.Server(lanes=3)
    bad, seen = [], [0]

    def check(i, k, c):
        seen[0] += 1
        if c["status"] != 0:
            bad.append((i, k, "status", c["status"]))
            return
        if _digest(*srv.read(c)) != alone[i][k % 8]:
            bad.append((i, k, "differs"))

    st = serve.run(srv, specs, fps=30.0, duration=0.6, seed=5, check=check)
    srv.close()
    _say("  stats", {k: st[k] for k in ("frames_offered", "frames_completed", "frames_failed", "batches",
                                         "frames_per_batch", "p50_ms", "p99_ms")})
    assert st["frames_completed"] == st["frames_offered"] == seen[0] > 64
    assert st["frames_failed"] == 0
    assert st["frames_per_batch"] > 1.0  # frames of several streams shared launches
    assert not bad, bad[:5]
