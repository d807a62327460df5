"""Quality metric: symmetric nearest-neighbour MSE (MSE::Compute,
registration_base.cc:168-223) and PSNR (pvvc_decompression.cc:156-159).

The oracle's grid search is pinned against its brute-force search and
against closed forms; the GPU metric is compared with the oracle: neighbour
indices exactly (ties go to the lowest index), FP64 sums to a relative
1e-9 (only the summation order differs)."""
import math

import numpy as np
import pytest

REL = 1e-9


def _tie_clouds(rng, n, m, half=False, scale=1.0):
    """Integer-grid clouds: many equidistant candidates."""
    p = rng.integers(0, 12, (n, 3)).astype(np.float32)
    q = (rng.integers(0, 12, (m, 3)) + (0.5 if half else 0.0)).astype(np.float32) * np.float32(scale)
    return (p, rng.integers(0, 256, (n, 3)).astype(np.uint8)), (q, rng.integers(0, 256, (m, 3)).astype(np.uint8))


def test_oracle_grid_equals_brute(oracle):
    rng = np.random.default_rng(3)
    for t in range(24):
        p, q = _tie_clouds(rng, int(rng.integers(1, 300)), int(rng.integers(1, 300)), half=t % 2 == 1,
                           scale=[1.0, 0.37, 2.9][t % 3])
        a = oracle.cloud_mse(*p, *q, mode=oracle.MSE_BRUTE, want_nn=True)
        b = oracle.cloud_mse(*p, *q, mode=0, want_nn=True)
        assert np.array_equal(a["nn_p"], b["nn_p"]) and np.array_equal(a["nn_q"], b["nn_q"]), t
        for k in ("geo", "y", "u", "v"):
            assert a[k] == b[k], (t, k)


def test_oracle_closed_forms(oracle):
    rng = np.random.default_rng(4)
    g = np.stack(np.meshgrid(*[np.arange(0, 30, 3)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    c = rng.integers(0, 256, g.shape).astype(np.uint8)
    same = oracle.cloud_mse(g, c, g, c)
    assert all(same[k] == (0.0, 0.0) for k in ("geo", "y", "u", "v"))
    # unit shift on a grid of spacing 3: every neighbour is the shifted copy
    sh = oracle.cloud_mse(g, c, g + np.float32([1, 0, 0]), c, mode=oracle.MSE_BRUTE)
    assert sh["geo"] == (1.0, 1.0) and sh["y"] == (0.0, 0.0)
    # grey offset of 10: Y differs by 10 (0.299 + 0.587 + 0.114 = 1); U, V by ~0
    g1 = np.full(g.shape, 100, np.uint8)
    g2 = np.full(g.shape, 110, np.uint8)
    grey = oracle.cloud_mse(g, g1, g, g2)
    assert abs(grey["y"][0] - 100.0) < 1e-3 and abs(grey["y"][1] - 100.0) < 1e-3
    assert grey["u"][0] < 1e-6 and grey["v"][0] < 1e-6
    # one point at the origin against three on the x axis
    o = np.zeros((1, 3), np.float32)
    x = np.array([[1, 0, 0], [2, 0, 0], [3, 0, 0]], np.float32)
    col1, col3 = np.zeros((1, 3), np.uint8), np.zeros((3, 3), np.uint8)
    r = oracle.cloud_mse(o, col1, x, col3)
    assert r["geo"][0] == 1.0 and abs(r["geo"][1] - 14.0 / 3.0) < 1e-12
    # subset: every p point has an exact match in q
    sub = oracle.cloud_mse(g[::2], c[::2], g, c)
    assert sub["geo"][0] == 0.0 and sub["y"][0] == 0.0 and sub["geo"][1] > 0.0
    # empty clouds (registration_base.cc:170-179)
    e = oracle.cloud_mse(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint8), g, c)
    assert e["geo"] == (-1.0, 0.0) and e["y"] == (-1.0, 0.0)
    e2 = oracle.cloud_mse(g, c, np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint8))
    assert e2["geo"] == (0.0, -1.0)


def test_oracle_float_running_sum_close(oracle):
    """The reference's float running sums stay within float accumulation
    error of the FP64 sums."""
    rng = np.random.default_rng(5)
    p, q = _tie_clouds(rng, 5000, 4000, half=True)
    a = oracle.cloud_mse(*p, *q)
    b = oracle.cloud_mse(*p, *q, mode=oracle.MSE_FLOAT_RUNNING)
    for k in ("geo", "y", "u", "v"):
        for i in (0, 1):
            assert abs(a[k][i] - b[k][i]) <= 1e-4 * abs(a[k][i]) + 1e-6


def test_psnr_formula(pvvc):
    """pvvc_psnr is host code: 10 log10(peak^2 / max(a, b)) in FP32."""
    v = pvvc.quality.psnr((0.5, 0.25), 1024.0)
    assert abs(v - 10 * math.log10(1024.0 ** 2 / 0.5)) < 1e-4
    assert abs(pvvc.quality.psnr((2.0, 4.0), 256.0) - 10 * math.log10(256.0 ** 2 / 4.0)) < 1e-4


def _check_gpu(pvvc, oracle, p, q, mode=0):
    g = pvvc.quality.cloud_mse(p, q, want_nn=True)
    o = oracle.cloud_mse(*p, *q, mode=mode, want_nn=True)
    assert np.array_equal(g["nn_p"], o["nn_p"]), np.flatnonzero(g["nn_p"] != o["nn_p"])[:5]
    assert np.array_equal(g["nn_q"], o["nn_q"]), np.flatnonzero(g["nn_q"] != o["nn_q"])[:5]
    for k in ("geo", "y", "u", "v"):
        for i in (0, 1):
            assert abs(g[k][i] - o[k][i]) <= REL * abs(o[k][i]) + 1e-12, (k, i, g[k][i], o[k][i])
    return g


@pytest.mark.gpu
def test_gpu_mse_small_ties(pvvc, oracle):
    rng = np.random.default_rng(6)
    for t in range(12):
        p, q = _tie_clouds(rng, int(rng.integers(1, 400)), int(rng.integers(1, 400)), half=t % 2 == 1,
                           scale=[1.0, 0.37, 2.9][t % 3])
        print(f"case {t}: n={len(p[0])} m={len(q[0])}", flush=True)
        _check_gpu(pvvc, oracle, p, q, mode=oracle.MSE_BRUTE)


@pytest.mark.gpu
def test_gpu_mse_outliers_and_degenerate(pvvc, oracle):
    rng = np.random.default_rng(7)
    # a dense blob and queries far outside it (past the shell limit: full scan)
    p = (rng.normal(0, 2, (3000, 3))).astype(np.float32)
    q = np.concatenate([rng.normal(0, 2, (200, 3)), rng.normal(0, 1, (20, 3)) + 900.0]).astype(np.float32)
    cp = rng.integers(0, 256, (len(p), 3)).astype(np.uint8)
    cq = rng.integers(0, 256, (len(q), 3)).astype(np.uint8)
    _check_gpu(pvvc, oracle, (p, cp), (q, cq), mode=oracle.MSE_BRUTE)
    # every point identical: one cell
    z = np.zeros((50, 3), np.float32)
    _check_gpu(pvvc, oracle, (z, cp[:50]), (z[:7] + 0.25, cq[:7]), mode=oracle.MSE_BRUTE)
    # single points
    _check_gpu(pvvc, oracle, (p[:1], cp[:1]), (q[:1], cq[:1]), mode=oracle.MSE_BRUTE)
    # empty direction semantics
    m = pvvc.quality.MSE()
    m.set_clouds((np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint8)), (p, cp))
    m.compute()
    assert m.get_geo_mses() == (-1.0, 0.0) and m.get_y_mses() == (-1.0, 0.0)


@pytest.mark.gpu
def test_gpu_mse_decoded_vs_source(pvvc, oracle):
    """A decoded c0 frame against the cloud it was encoded from, full size:
    exact neighbours, FP64 sums to 1e-9; PSNR as the reference prints it."""
    frames, src = pvvc.synth_stream("c0", n_frames=2, seed=7, threads=8, keep_source=True)
    dec = pvvc.Decoder(max_patches=4096, max_points=8192)
    fr = [pvvc.Frame.parse(b) for b in frames]
    batch = dec.batch(fr)
    xyz, rgb = dec.decode_host(batch)
    assert (batch.status() == 0).all()
    for f in range(2):
        p = src[f]
        n = batch.frame_points[f]  # voxels that share a leaf decode as one point
        assert 0.99 * len(p[0]) <= n <= len(p[0])
        off = batch.frame_offset[f]
        d = (xyz[off:off + n], rgb[off:off + n])
        g = _check_gpu(pvvc, oracle, p, d)
        geo_db = pvvc.quality.psnr(g["geo"], pvvc.quality.GEO_PEAK)
        y_db = pvvc.quality.psnr(g["y"], pvvc.quality.Y_PEAK)
        print(f"frame {f}: {n} pts, GEO {geo_db:.2f} dB, Y {y_db:.2f} dB, {g['kernel_ms']:.2f} ms", flush=True)
        assert 20.0 < y_db < 80.0 and geo_db > 40.0


@pytest.mark.gpu
def test_gpu_mse_million_points(pvvc, oracle):
    """1M-point voxel surface against a jittered copy: grid oracle vs GPU."""
    rng = np.random.default_rng(8)
    n = 1 << 20
    th = rng.uniform(0, 2 * np.pi, n)
    ph = rng.uniform(0, np.pi, n)
    r = 400.0
    p = np.floor(np.stack([512 + r * np.sin(ph) * np.cos(th), 512 + r * np.sin(ph) * np.sin(th),
                           512 + r * np.cos(ph)], -1)).astype(np.float32)
    q = p + rng.integers(-1, 2, p.shape).astype(np.float32)
    cp = rng.integers(0, 256, p.shape).astype(np.uint8)
    cq = np.clip(cp.astype(np.int32) + rng.integers(-3, 4, p.shape), 0, 255).astype(np.uint8)
    g = _check_gpu(pvvc, oracle, (p, cp), (q, cq))
    print(f"1M: {g['kernel_ms']:.2f} ms, geo {g['geo']}", flush=True)
