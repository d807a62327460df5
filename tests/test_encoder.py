"""Forward path: the GPU colour encoder (forward RAHT, quantisation, RLGR
writer) against the oracle's restatement of the encoder's colour half
(oracle/pvvc_oracle.cpp orc_color_encode: raht_octree.cc:189-247,
GoP_encoding.cc:172-192, entropy_codec.cc:65-159), byte for byte, on
surface-like voxel patches with intra-like and residual-like colours and on
every colour stream of a decoded c1 frame.  The stream writer's CPU encoder
(product code, the synthetic-stream generator) only builds the inputs here,
and is itself checked against the oracle."""
import sys

import numpy as np
import pytest


def _surface_patch(rng, n_target):
    """Unique voxels near a random sphere cap: a dense surface like a patch."""
    r = rng.uniform(8, 40)
    pts = set()
    c = rng.uniform(50, 200, 3)
    tries = 0
    while len(pts) < n_target and tries < 20 * n_target:
        tries += 1
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        if v[2] < 0.2:
            continue
        p = np.floor(c + r * v + rng.normal(0, 0.3, 3))
        pts.add((p[0], p[1], p[2]))
    return np.array(sorted(pts), np.float32)


def _jobs(pvvc, oracle, seed=3, count=40):
    """(octree bytes, leaf YUV, qp) jobs and the oracle's bytes for each; the
    writer's own bytes must equal them too."""
    rng = np.random.default_rng(seed)
    jobs, refs = [], []
    sizes = [1, 2, 3, 9] + list(rng.integers(10, 3000, count - 4))
    for t, n in enumerate(sizes):
        xyz = _surface_patch(rng, int(n)) if n > 3 else np.array([[5, 7, 9], [6, 7, 9], [5, 8, 9]][: n],
                                                                    np.float32)
        m = len(xyz)
        base = 128 + 60 * np.sin(xyz @ rng.normal(0, 0.05, (3, 3)))
        yuv = (base + rng.normal(0, 4, (m, 3))).astype(np.float32)
        if t % 3 == 2:  # residual-like: signed, small
            yuv = rng.normal(0, 6, (m, 3)).astype(np.float32)
        qp = int(rng.choice([1, 2, 4, 8]))
        geo, leaf, col = pvvc.encode_colors_cpu(xyz, yuv, qp)
        # leaves may hold several points (the tree's range is max - min, so
        # neighbouring voxels can share a leaf); the encoders take leaf colours
        assert 1 <= len(leaf) <= m
        ref, _, st = oracle.color_encode(geo, leaf, qp)
        assert st == 0 and ref == col
        jobs.append((geo, leaf, qp))
        refs.append(ref)
    return jobs, refs


def test_cpu_encoder_decodes_with_oracle(pvvc, oracle):
    """The CPU reference encoder's stream decodes (oracle) to the RAHT of the
    leaf colours within the quantisation step."""
    jobs, refs = _jobs(pvvc, oracle, seed=5, count=8)
    for (geo, leaf, qp), col in zip(jobs, refs):
        p = oracle.Patch()
        assert p.set_slice(0x01, len(leaf), qp, geo, col) == 0
        out = p.get(len(leaf))
        # RAHT is orthonormal: per-leaf error is bounded by the per-coefficient error
        err = np.abs(out["yuv"].astype(np.float64) - leaf.astype(np.float64)).max()
        assert err <= qp * np.sqrt(3 * len(leaf)) / 2 + 1e-3


@pytest.mark.gpu
def test_gpu_encoder_matches_oracle(pvvc, oracle):
    jobs, refs = _jobs(pvvc, oracle)
    outs, status, ms = pvvc.encode_colors(jobs)
    assert status == [0] * len(jobs), status
    for i, (o, r) in enumerate(zip(outs, refs)):
        assert o == r, f"job {i}: n={len(jobs[i][1])} qp={jobs[i][2]} {len(o)} vs {len(r)} bytes"
    assert ms > 0


@pytest.mark.gpu
def test_gpu_encoder_rejects_bad_geometry(pvvc, oracle):
    geo, leaf, qp = _jobs(pvvc, oracle, seed=9, count=5)[0][4]
    bad = bytearray(geo)
    bad[24] = 0  # height 0
    outs, status, _ = pvvc.encode_colors([(bytes(bad), leaf, qp), (geo, leaf[:-1], qp)])
    assert status == [1, 3]  # BAD_SLICE, UNMATCHED_SIZE (leaf count != n)


def test_cpu_encoder_on_tree_equals_point_encoder(pvvc, oracle):
    """The writer's encoder on given octree bytes produces the same stream as
    the writer's encoder from points (and as the oracle)."""
    jobs, refs = _jobs(pvvc, oracle, seed=11, count=12)
    for (geo, leaf, qp), col in zip(jobs, refs):
        assert pvvc.encode_colors_cpu_on_tree(geo, leaf, qp) == col


def _boundary_jobs(pvvc, oracle):
    """Streams whose byte length ends on or just short of a 256-byte (64-word)
    boundary: the writer's last partial word then completes or nearly
    completes a store chunk.  Found by a seeded search over small random
    patches; every residue class below is required."""
    rng = np.random.default_rng(21)
    want = {253: None, 254: None, 255: None, 0: None, 1: None, 252: None}
    for _ in range(4000):
        n = int(rng.integers(40, 700))
        v = np.unique(rng.integers(0, 24, (n, 3)), axis=0).astype(np.float32)
        yuv = (128 + rng.normal(0, rng.uniform(2, 30), (len(v), 3))).astype(np.float32)
        qp = int(rng.choice([1, 2, 4]))
        geo, leaf, col = pvvc.encode_colors_cpu(v, yuv, qp)
        r = len(col) % 256
        if len(col) >= 256 and r in want and want[r] is None:
            ref, _, st = oracle.color_encode(geo, leaf, qp)
            assert st == 0 and ref == col
            want[r] = ((geo, leaf, qp), ref)
        if all(x is not None for x in want.values()):
            break
    assert all(x is not None for x in want.values()), {k: x is not None for k, x in want.items()}
    return [x[0] for x in want.values()], [x[1] for x in want.values()]


def test_boundary_search_finds_every_residue(pvvc, oracle):
    jobs, refs = _boundary_jobs(pvvc, oracle)
    assert sorted(len(r) % 256 for r in refs) == [0, 1, 252, 253, 254, 255]


@pytest.mark.gpu
def test_gpu_encoder_chunk_boundaries(pvvc, oracle):
    jobs, refs = _boundary_jobs(pvvc, oracle)
    outs, status, _ = pvvc.encode_colors(jobs)
    assert status == [0] * len(jobs), status
    for o, r in zip(outs, refs):
        assert o == r, f"{len(r)} bytes (mod 256 = {len(r) % 256})"


def test_packed_encoder_rejects_inconsistent_jobs(pvvc):
    """Argument checks of the packed entry (host side, before any GPU call)."""
    geo = np.zeros(30, np.uint8)
    yuv = np.zeros((4, 3), np.float32)
    with pytest.raises(pvvc.PvvcError):
        pvvc.encode_colors_packed(geo, [30], yuv, [4, 1], np.array([1], np.uint8))
    with pytest.raises(pvvc.PvvcError):
        pvvc.encode_colors_packed(geo, [31], yuv, [4], np.array([1], np.uint8))
    with pytest.raises(pvvc.PvvcError):
        pvvc.encode_colors_packed(geo, [30], yuv, [5], np.array([1], np.uint8))


@pytest.mark.gpu
def test_gpu_encoder_kept_buffers(pvvc, oracle):
    """Calls in a row reuse the library's device and pinned buffers: a large
    call (~1.2 M leaves, ~14 MB of input, several staging chunks), a smaller
    one, the large one again in reverse order, then again after the buffers
    are released.  Every stream equals the oracle's bytes."""
    jobs, refs = _jobs(pvvc, oracle, seed=31, count=40)
    reps = 20
    big, big_refs = jobs * reps, refs * reps
    assert sum(len(j[1]) for j in big) * 12 > (8 << 20)
    small, small_refs = _boundary_jobs(pvvc, oracle)
    for name, js, rs in [("big", big, big_refs), ("small", small, small_refs),
                         ("big reversed", big[::-1], big_refs[::-1]), ("small", small, small_refs)]:
        outs, status, _ = pvvc.encode_colors(js)
        assert status == [0] * len(js), name
        bad = [i for i, (o, r) in enumerate(zip(outs, rs)) if o != r]
        assert not bad, f"{name}: {len(bad)} streams differ, first {bad[:5]}"
    pvvc.hip_lib().pvvc_encoder_release()
    outs, status, _ = pvvc.encode_colors(small)
    assert status == [0] * len(small) and outs == small_refs


@pytest.mark.gpu
def test_gpu_packed_encoder_equals_job_list(pvvc, oracle):
    jobs, refs = _jobs(pvvc, oracle, seed=41, count=24)
    geo_all = np.frombuffer(b"".join(bytes(j[0]) for j in jobs), np.uint8)
    gsz = [len(j[0]) for j in jobs]
    yuv_all = np.concatenate([j[1] for j in jobs])
    npts = [len(j[1]) for j in jobs]
    qp = np.array([j[2] for j in jobs], np.uint8)
    out_all, off, size, status, ms = pvvc.encode_colors_packed(geo_all, gsz, yuv_all, npts, qp)
    assert (status == 0).all() and ms > 0
    assert [out_all[o:o + z].tobytes() for o, z in zip(off, size)] == refs


@pytest.mark.gpu
def test_gpu_encoder_c1_frame_streams_match_oracle(pvvc, oracle):
    """Every colour stream of a c1 intra frame (~860 patches of 200-3000
    points): the octree bytes of each slice, its leaf colours as the oracle
    decodes them, and its qp, encoded on the GPU and by the oracle; plus a
    signed residual-like colour set on each of the same trees."""
    imgs = pvvc.synth_stream("c1", n_frames=1, seed=17, threads=16)
    f0 = pvvc.Frame.parse(imgs[0])
    m0 = f0.meta()
    jobs, want = [], []
    for s in range(f0.slice_cnt):
        geo = f0.geometry(s)
        n, qp = int(m0["size"][s]), int(m0["qp"][s])
        pt = oracle.Patch()
        assert pt.set_slice(0x01, n, qp, geo, f0.color(s)) == 0
        leaf = pt.get(n)["yuv"]
        for yuv in (leaf, leaf - leaf[::-1]):
            jobs.append((geo, yuv, qp))
            ref, _, st = oracle.color_encode(geo, yuv, qp)
            assert st == 0
            want.append(ref)
    print(f"c1 frame 0: {f0.slice_cnt} slices, {len(jobs)} colour streams", flush=True)
    sys.stdout.flush()
    outs, status, ms = pvvc.encode_colors(jobs)
    assert status == [0] * len(jobs)
    bad = [i for i, (o, r) in enumerate(zip(outs, want)) if o != r]
    assert not bad, f"{len(bad)} of {len(jobs)} streams differ, first {bad[:5]}"
