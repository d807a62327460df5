"""Pins the CPU oracle (oracle/pvvc_oracle.cpp) before it is trusted.

The reference ships no golden vectors for this path, so each part is pinned
by checks that a plausible restatement mistake fails:

* RLGR: known-answer vectors derived by hand from entropy_codec.cc:65-159
  (derivations inline), plus encode->decode round trips over adversarial
  symbol distributions.
* Octree: closed-form leaf centres for hand-built geometry headers, and a
  from-scratch recursive expansion written here.
* Inverse RAHT: the two-leaf closed form, the constant-colour identity
  (DC = sqrt(n) * c, AC = 0) and forward->inverse identity, against FP64.
* Forward RAHT + quantisation (the GPU encoder's oracle): the two-leaf
  closed form in exact FP32, the constant-colour identity, identity through
  the pinned inverse with Parseval's equality, and round-half-away-from-zero
  known answers on single-leaf patches.
* ConvertRGB: BT.601 inverse matrix values and rounding at .5 boundaries.
"""
import math
import struct

import numpy as np
import pytest

# --------------------------------------------------------------------------
# RLGR known answers.  State starts k_P = 0 (k = 0), k_RP = 8 (k_R = 2).
# --------------------------------------------------------------------------
KATS = [
    # [1]: u = 2, k = 0 -> GR(2, k_R=2): quotient 0 -> "0", remainder "10"
    #      -> "010" padded -> 0x40
    ([1], "40"),
    # [0]: u = 0 -> "0" + "00" -> 0x00
    ([0], "00"),
    # [-1]: u = 1 -> "0" + "01" -> 0x20
    ([-1], "20"),
    # [0,0,0,0,0]: s0 "000" (k_RP 8->6, k_P 0->3); s1 k_R=1: "00" (k_RP->4,
    # k_P->6); s2 k=1 run, m=1; s3 m=2=2^1 -> "1", k_P=8; s4 k=2, m=1;
    # end: "0" + m=1 in k_P/4=2 bits "01" -> 000 00 1 0 01 -> 04 80
    ([0, 0, 0, 0, 0], "0480"),
    # [0,0,0,5]: "000","00", s2 run m=1, s3 u=10 -> u-1=9: "0" + m "1" +
    # GR(9, k_R=1) = "11110"+"1" -> 000000 1 111101 -> 03 E8
    ([0, 0, 0, 5], "03e8"),
    # [200]: u = 400, 400>>2 = 100 >= 32 -> 32 ones then 400 in 32 bits
    ([200], "ffffffff00000190"),
]


@pytest.mark.parametrize("sym,hexbytes", KATS)
def test_rlgr_known_answers(oracle, sym, hexbytes):
    assert oracle.rlgr_encode(sym).hex() == hexbytes
    dec, st = oracle.rlgr_decode(bytes.fromhex(hexbytes), len(sym))
    assert st == 0
    assert dec.tolist() == sym


def _random_symbols(rng, n):
    kind = rng.integers(0, 5)
    if kind == 0:   # sparse, long zero runs
        v = rng.integers(-3, 4, n) * (rng.random(n) < 0.05)
    elif kind == 1:  # dense small
        v = rng.integers(-2, 3, n)
    elif kind == 2:  # large values near the int32 limit (escape codes)
        v = rng.integers(-(2**31) + 1, 2**31 - 1, n)
    elif kind == 3:  # mixture: runs of zeros broken by big spikes
        v = np.where(rng.random(n) < 0.02, rng.integers(-70000, 70000, n), 0)
    else:           # geometric-ish magnitudes
        v = (rng.geometric(0.2, n) - 1) * rng.choice([-1, 1], n)
    return v.astype(np.int64)


def test_rlgr_round_trip(oracle):
    rng = np.random.default_rng(1234)
    for t in range(300):
        n = int(rng.choice([1, 2, 3, 7, 64, 65, 1000, 4096, 100000 if t % 50 == 0 else 5000]))
        v = _random_symbols(rng, n)
        b = oracle.rlgr_encode(v)
        d, st = oracle.rlgr_decode(b, n)
        assert st == 0, (t, n)
        np.testing.assert_array_equal(d, v)


def test_rlgr_truncation_flags_overrun(oracle):
    rng = np.random.default_rng(5)
    v = rng.integers(-20, 20, 3000)
    b = oracle.rlgr_encode(v)
    _, st = oracle.rlgr_decode(b[: len(b) // 2], 3000)
    assert st == 5  # BITSTREAM
    _, st = oracle.rlgr_decode(b"", 10)
    assert st == 5


def test_product_rlgr_encoder_matches_oracle(oracle, pvvc):
    """The stream writer's RLGR encoder and the oracle's agree byte for byte."""
    L = pvvc.enc_lib()
    rng = np.random.default_rng(99)
    for t in range(100):
        v = _random_symbols(rng, int(rng.integers(1, 20000)))
        n = L.pvvc_enc_rlgr(v.ctypes.data, v.size, None, 0)
        out = np.zeros(n, np.uint8)
        L.pvvc_enc_rlgr(v.ctypes.data, v.size, out.ctypes.data, n)
        assert out.tobytes() == oracle.rlgr_encode(v)


# --------------------------------------------------------------------------
# Octree
# --------------------------------------------------------------------------
def header(center, rng, h):
    """SaveTreeCore layout: 6 big-endian f32 then u8 height (octree_base.cc:20-31)."""
    return b"".join(struct.pack(">f", x) for x in (*center, *rng)) + bytes([h])


def test_octree_header_bytes():
    # centre (1,1,1) -> 3F800000 x3, range (2,2,2) -> 40000000 x3, h=2 -> 02
    assert header((1, 1, 1), (2, 2, 2), 2).hex() == "3f800000" * 3 + "40000000" * 3 + "02"


def test_octree_full_node(oracle):
    # children 0..7: x/y/z "> centre ? 0 : 1" per bit (raht_octree.cc:141-148);
    # child offset = (range/2)/2 = 0.5
    xyz, st, h = oracle.tree_expand(header((1, 1, 1), (2, 2, 2), 2) + b"\xff", 8)
    assert st == 0 and h == 2
    expect = [(1.5 if not (i & 4) else 0.5, 1.5 if not (i & 2) else 0.5, 1.5 if not (i & 1) else 0.5)
              for i in range(8)]
    np.testing.assert_array_equal(xyz, np.array(expect, np.float32))


def test_octree_two_corners(oracle):
    xyz, st, _ = oracle.tree_expand(header((1, 1, 1), (2, 2, 2), 2) + b"\x81", 2)
    assert st == 0
    np.testing.assert_array_equal(xyz, np.array([(1.5, 1.5, 1.5), (0.5, 0.5, 0.5)], np.float32))


def test_octree_single_leaf(oracle):
    xyz, st, _ = oracle.tree_expand(header((3.5, -2, 7), (1, 1, 1), 1), 1)
    assert st == 0
    np.testing.assert_array_equal(xyz, np.array([(3.5, -2, 7)], np.float32))


def test_octree_errors(oracle):
    _, st, _ = oracle.tree_expand(header((0, 0, 0), (4, 4, 4), 3) + b"\x81", 2)  # truncated level 1
    assert st == 1
    _, st, _ = oracle.tree_expand(header((0, 0, 0), (4, 4, 4), 2) + b"\x81", 3)  # leaf count != size
    assert st == 3
    _, st, _ = oracle.tree_expand(header((0, 0, 0), (4, 4, 4), 0), 1)  # height 0
    assert st == 1


def _expand_recursive(geo: bytes):
    """From-scratch expansion: read levels breadth-first, then walk depth-first
    computing centres as centre +- (range/2)/2 per set bit, in float32."""
    c = np.frombuffer(geo[:24], ">f4").astype(np.float32)
    h = geo[24]
    levels, pos, cnt = [], 25, 1
    for _ in range(h - 1):
        lv = list(geo[pos:pos + cnt])
        pos += cnt
        levels.append(lv)
        cnt = sum(bin(b).count("1") for b in lv)
    out = []

    def walk(depth, idx, cen, rng):
        if depth == h - 1:
            out.append(cen)
            return
        b = levels[depth][idx]
        first = sum(bin(x).count("1") for x in levels[depth][:idx])
        sub = (rng / np.float32(2)).astype(np.float32)
        k = 0
        for i in range(8):
            if b & (0x80 >> i):
                off = (sub / np.float32(2)).astype(np.float32)
                sgn = np.array([-1 if i & 4 else 1, -1 if i & 2 else 1, -1 if i & 1 else 1], np.float32)
                walk(depth + 1, first + k, (cen + sgn * off).astype(np.float32), sub)
                k += 1

    walk(0, 0, c[:3], c[3:])
    return np.array(out, np.float32)


def _random_geometry(rng, h, density):
    lv, geo, cnt = [], b"", 1
    for _ in range(h - 1):
        bs = []
        for _ in range(cnt):
            b = 0
            while b == 0:
                b = int(sum((0x80 >> i) for i in range(8) if rng.random() < density))
            bs.append(b)
        geo += bytes(bs)
        cnt = sum(bin(b).count("1") for b in bs)
    cen = rng.uniform(-100, 100, 3).round(1)
    r = 2.0 ** (h - 1)
    return header(tuple(cen), (r, r, r), h) + geo, cnt


def test_octree_against_recursive_expansion(oracle):
    rng = np.random.default_rng(3)
    for t in range(40):
        h = int(rng.integers(1, 8))
        geo, n = _random_geometry(rng, h, float(rng.uniform(0.15, 0.6)))
        xyz, st, _ = oracle.tree_expand(geo, n)
        assert st == 0
        np.testing.assert_array_equal(xyz, _expand_recursive(geo))


def test_encoder_geometry_round_trip(oracle, pvvc):
    """Stream-writer octree -> oracle expansion: every input voxel lies in the
    unit cell of a decoded leaf and every leaf holds at least one voxel."""
    L = pvvc.enc_lib()
    rng = np.random.default_rng(11)
    import ctypes as C
    for t in range(20):
        n = int(rng.integers(1, 3000))
        ext = int(rng.integers(1, 200))
        pts = (np.unique(rng.integers(0, ext, (n, 3)), axis=0) + rng.integers(0, 500, 3)).astype(np.float32)
        leaves = C.c_int()
        sz = L.pvvc_enc_octree(pts.ctypes.data, len(pts), None, 0, C.byref(leaves))
        buf = np.zeros(sz, np.uint8)
        L.pvvc_enc_octree(pts.ctypes.data, len(pts), buf.ctypes.data, sz, C.byref(leaves))
        xyz, st, h = oracle.tree_expand(buf.tobytes(), leaves.value)
        assert st == 0
        assert len(np.unique(xyz, axis=0)) == len(xyz)
        d = np.abs(pts[:, None, :] - xyz[None, :, :]).max(-1) if len(pts) * len(xyz) < 4e6 else None
        if d is not None:
            assert (d.min(1) <= 0.5).all()
            assert (d.min(0) <= 0.5).all()


# --------------------------------------------------------------------------
# Inverse RAHT
# --------------------------------------------------------------------------
def _slice_colour_bytes(oracle, y, u, v):
    return oracle.rlgr_encode(np.concatenate([y, u, v]).astype(np.int64))


def test_iraht_two_leaf_closed_form(oracle):
    """occupancy 0x81: one coded AC at idx 1 (w2 = w3 = 1); stream order is
    [AC, DC]; leaves are (DC - h)/sqrt2 and (DC + h)/sqrt2."""
    geo = header((1, 1, 1), (2, 2, 2), 2) + b"\x81"
    dc = np.array([300, 200, 180])
    hh = np.array([-40, 12, 7])
    col = _slice_colour_bytes(oracle, [hh[0], dc[0]], [hh[1], dc[1]], [hh[2], dc[2]])
    p = oracle.Patch()
    assert p.set_slice(0x01, 2, 1, geo, col) == 0
    r = p.get(2)
    exp = np.stack([(dc - hh) / math.sqrt(2), (dc + hh) / math.sqrt(2)])
    np.testing.assert_allclose(r["yuv"], exp, rtol=0, atol=1e-4)
    # the FP32 result equals the reference op order g*a + h*(-c)
    a = np.float32(np.sqrt(np.float32(1)) / np.sqrt(np.float32(2)))
    lo = np.float32(dc[0]) * a + np.float32(hh[0]) * (-a)
    assert r["yuv"][0, 0] == lo


def test_iraht_constant_colour(oracle, pvvc):
    """Forward RAHT of a constant colour: DC = sqrt(n) * c, all ACs 0; the
    decoded patch is the constant again."""
    import ctypes as C
    L = pvvc.enc_lib()
    rng = np.random.default_rng(2)
    pts = np.unique(rng.integers(0, 40, (700, 3)), axis=0).astype(np.float32)
    n = len(pts)
    yuv = np.tile(np.array([[100.0, 120.0, 140.0]], np.float32), (n, 1))
    coef = np.zeros((n, 3), np.float32)
    k = L.pvvc_enc_raht(pts.ctypes.data, n, yuv.ctypes.data, coef.ctypes.data)
    assert k == n
    np.testing.assert_allclose(coef[-1], math.sqrt(n) * yuv[0], rtol=1e-5)
    np.testing.assert_allclose(coef[:-1], 0, atol=2e-3)
    # decode a slice carrying exactly those (integer) coefficients
    q = np.rint(coef).astype(np.int64)
    leaves = C.c_int()
    sz = L.pvvc_enc_octree(pts.ctypes.data, n, None, 0, C.byref(leaves))
    g = np.zeros(sz, np.uint8)
    L.pvvc_enc_octree(pts.ctypes.data, n, g.ctypes.data, sz, C.byref(leaves))
    p = oracle.Patch()
    assert p.set_slice(0x01, n, 1, g.tobytes(), _slice_colour_bytes(oracle, q[:, 0], q[:, 1], q[:, 2])) == 0
    r = p.get(n)
    np.testing.assert_allclose(r["rgb64"], r["rgb64"][0:1].repeat(n, 0), atol=1e-9)
    exp = q[-1] / math.sqrt(n)
    np.testing.assert_allclose(r["yuv"], np.tile(exp, (n, 1)), atol=2e-3)


def test_iraht_forward_inverse_identity(oracle, pvvc):
    """Forward (stream writer) then inverse (oracle) with qp = 1 returns the
    colours up to the rounding of the coefficients (orthonormal transform:
    the per-leaf error stays within a few quantisation steps)."""
    import ctypes as C
    L = pvvc.enc_lib()
    rng = np.random.default_rng(8)
    for t in range(6):
        pts = np.unique(rng.integers(0, 64, (int(rng.integers(50, 2500)), 3)), axis=0).astype(np.float32)
        n = len(pts)
        yuv = rng.uniform(0, 255, (n, 3)).astype(np.float32)
        coef = np.zeros((n, 3), np.float32)
        assert L.pvvc_enc_raht(pts.ctypes.data, n, yuv.ctypes.data, coef.ctypes.data) == n
        q = np.rint(coef).astype(np.int64)
        leaves = C.c_int()
        sz = L.pvvc_enc_octree(pts.ctypes.data, n, None, 0, C.byref(leaves))
        g = np.zeros(sz, np.uint8)
        L.pvvc_enc_octree(pts.ctypes.data, n, g.ctypes.data, sz, C.byref(leaves))
        p = oracle.Patch()
        assert p.set_slice(0x01, n, 1, g.tobytes(), _slice_colour_bytes(oracle, q[:, 0], q[:, 1], q[:, 2])) == 0
        r = p.get(n)
        # map decoded leaves back to input points (leaf centre within 0.5)
        xyz = r["xyz"]
        order = np.lexsort(xyz.T[::-1])
        pin = np.lexsort((pts + 0.5).T[::-1])
        err = np.abs(r["yuv"][order] - yuv[pin])
        assert np.sqrt((err ** 2).mean()) < 0.6
        assert err.max() < 4.0


# --------------------------------------------------------------------------
# Forward RAHT and quantisation (the oracle of the GPU colour encoder)
# --------------------------------------------------------------------------
def test_forward_two_leaf_closed_form(oracle):
    """Occupancy 81 (children 0 and 7 of one node): every pair but the root's
    passes one side through (factors 0 and 1), and the root pair has weights
    (1, 1), so the stream is [AC, DC] = [a*(-t) + b*t, a*t + b*t] with
    t = sqrt(1)/sqrt(2) in FP32 (HaarTransform, octree_base.cc:164-184;
    the final std::reverse, raht_octree.cc:245)."""
    geo = header((1, 1, 1), (2, 2, 2), 2) + b"\x81"
    a = np.array([10.25, -3.5, 200.0], np.float32)
    b = np.array([7.0, 91.75, -0.5], np.float32)
    c, st = oracle.raht_forward(geo, np.stack([a, b]))
    assert st == 0
    t = np.float32(np.sqrt(np.float32(1))) / np.sqrt(np.float32(2))
    np.testing.assert_array_equal(c[:, 0], a * (-t) + b * t)
    np.testing.assert_array_equal(c[:, 1], a * t + b * t)
    np.testing.assert_allclose(c[:, 1], (a.astype(np.float64) + b) / math.sqrt(2), rtol=1e-6)


def test_forward_constant_colour(oracle):
    """A constant colour c over n leaves: DC = sqrt(n) * c, every AC 0."""
    rng = np.random.default_rng(4)
    for t in range(12):
        geo, n = _random_geometry(rng, int(rng.integers(2, 8)), float(rng.uniform(0.2, 0.7)))
        col = np.array([100.0, 120.0, 140.0], np.float32)
        c, st = oracle.raht_forward(geo, np.tile(col, (n, 1)))
        assert st == 0
        np.testing.assert_allclose(c[:, -1], math.sqrt(n) * col, rtol=2e-6)
        np.testing.assert_allclose(c[:, :-1], 0, atol=1e-3 * math.sqrt(n))


def test_forward_inverse_identity(oracle):
    """The oracle's inverse (pinned above by its own closed forms) undoes the
    forward transform on random trees to FP32 accuracy: a wrong sign, a
    transposed factor, a wrong weight pair or coefficient order fails."""
    rng = np.random.default_rng(6)
    for t in range(30):
        geo, n = _random_geometry(rng, int(rng.integers(1, 9)), float(rng.uniform(0.1, 0.8)))
        leaf = rng.uniform(-50, 300, (n, 3)).astype(np.float32)
        c, st = oracle.raht_forward(geo, leaf)
        assert st == 0
        back, st = oracle.raht_inverse(geo, c)
        assert st == 0
        np.testing.assert_allclose(back, leaf, atol=2e-3 * max(1.0, math.log2(n + 1)))
        # Parseval: the transform is orthonormal
        np.testing.assert_allclose((c.astype(np.float64) ** 2).sum(), (leaf.astype(np.float64) ** 2).sum(),
                                   rtol=1e-5)


def test_forward_errors(oracle):
    geo = header((1, 1, 1), (2, 2, 2), 2) + b"\x81"
    assert oracle.raht_forward(geo, np.zeros((3, 3), np.float32))[1] == 3  # leaf count != n
    assert oracle.raht_forward(geo[:20], np.zeros((2, 3), np.float32))[1] == 1  # short header
    assert oracle.color_encode(geo, np.zeros((2, 3), np.float32), 0)[2] == 1  # qp 0


def test_quantisation_rounds_half_away_from_zero(oracle):
    """A single-leaf patch (height 1) codes its colour as the DC; with qp 4,
    10/4 = 2.5 -> 3, -10/4 -> -3, 6/4 = 1.5 -> 2 (std::round on c / QP,
    GoP_encoding.cc:180-186), and the bytes are the RLGR of those symbols."""
    geo = header((0, 0, 0), (1, 1, 1), 1)
    col, q, st = oracle.color_encode(geo, np.array([[10.0, -10.0, 6.0]], np.float32), 4)
    assert st == 0
    assert q.tolist() == [3, -3, 2]
    assert col == oracle.rlgr_encode([3, -3, 2])
    col, q, st = oracle.color_encode(geo, np.array([[9.99, -9.99, 5.99]], np.float32), 4)
    assert q.tolist() == [2, -2, 1]


def test_colour_stream_round_trip(oracle):
    """Encode then decode (both oracle): leaf colours come back within the
    quantisation bound of an orthonormal transform."""
    rng = np.random.default_rng(9)
    for t in range(10):
        geo, n = _random_geometry(rng, int(rng.integers(2, 7)), float(rng.uniform(0.2, 0.7)))
        leaf = rng.uniform(0, 255, (n, 3)).astype(np.float32)
        qp = int(rng.choice([1, 2, 4, 8]))
        col, q, st = oracle.color_encode(geo, leaf, qp)
        assert st == 0
        p = oracle.Patch()
        assert p.set_slice(0x01, n, qp, geo, col) == 0
        out = p.get(n)
        err = np.abs(out["yuv"].astype(np.float64) - leaf)
        assert err.max() <= qp * math.sqrt(3 * n) / 2 + 1e-2


def test_product_writer_matches_oracle_forward(oracle, pvvc):
    """The product's CPU stream writer (the synthetic-stream generator) writes
    the oracle's colour bytes on the same tree and leaf colours."""
    rng = np.random.default_rng(12)
    for t in range(25):
        geo, n = _random_geometry(rng, int(rng.integers(1, 8)), float(rng.uniform(0.1, 0.8)))
        leaf = (128 + rng.normal(0, rng.uniform(1, 60), (n, 3))).astype(np.float32)
        if t % 3 == 2:
            leaf = rng.normal(0, 8, (n, 3)).astype(np.float32)
        qp = int(rng.choice([1, 2, 4, 8, 16]))
        col, _, st = oracle.color_encode(geo, leaf, qp)
        assert st == 0
        assert pvvc.encode_colors_cpu_on_tree(geo, leaf, qp) == col, (t, n, qp)


# --------------------------------------------------------------------------
# ConvertRGB and motion
# --------------------------------------------------------------------------
def test_convert_rgb_values(oracle):
    """Grey (Y, 128, 128) maps to Y in all channels; a single-point patch with
    DC = (y, u, v) checks each inverse-BT.601 coefficient and the
    round-half-away-from-zero + saturation rule."""
    geo = header((0, 0, 0), (1, 1, 1), 1)
    for yuv, exp in [((100, 128, 128), (100, 100, 100)),
                     ((0, 128, 228), (140, 0, 0)),        # r = 1.402*100 = 140.2
                     ((100, 28, 128), (100, 134, 0)),     # g = 100+34.41, b = 100-177.2 -> 0
                     ((250, 228, 228), (255, 144, 255))]: # saturation
        p = oracle.Patch()
        col = oracle.rlgr_encode(np.array(yuv, np.int64))
        assert p.set_slice(0x01, 1, 1, geo, col) == 0
        r = p.get(1)
        assert tuple(r["rgb"][0]) == exp, (yuv, r["rgb"][0], r["rgbf"][0])
