"""RLGR stress: crafted symbol streams through every path of the GPU
decoder's scalar loop, against the oracle.

Intra slices over real octrees carry colour streams that the oracle's RLGR
encoder makes from symbol vectors built to reach:

* long zero runs: run mode with k up to ~20, runs crossing 64-symbol chunks,
  run prefixes of one to many one-bits (past the loop's 6, into the general
  path);
* Golomb escapes (32 one-bits and a raw 32-bit value) and codes longer than
  32 bits (large k_R after large values), in run mode and in k = 0 mode;
* frequent k = 0 / run mode changes, zero and non-zero k = 0 symbols;
* short runs in the k = 1 and k = 2 run states that reach chunk ends and, at
  the slice's end, the end of the slice;
* plain small random values.

qp is 1, so every coefficient difference shows in the unclamped drop-in
colours (float / 255), compared bit for bit with the oracle's."""
import sys

import numpy as np
import pytest

from _frames import pack_frame

IDENT = [1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _symbols(kind, L, rng):
    if kind == "sparse":
        s = np.zeros(L, np.int64)
        m = rng.random(L) < 0.02
        s[m] = rng.integers(1, 50, m.sum()) * rng.choice([-1, 1], m.sum())
        return s
    if kind == "bursty":
        s = np.zeros(L, np.int64)
        for b in range(0, L, 120):
            e = min(L, b + 20)
            s[b:e] = rng.integers(-9, 10, e - b)
        return s
    if kind == "huge":
        return rng.integers(-(1 << 20), 1 << 20, L)
    if kind == "mixed":
        s = np.zeros(L, np.int64)
        r = rng.random(L)
        small = (r > 0.70) & (r <= 0.95)
        large = r > 0.95
        s[small] = rng.integers(-6, 7, small.sum())
        s[large] = rng.integers(-(1 << 16), 1 << 16, large.sum())
        return s
    if kind == "ramp":
        a = np.arange(L, dtype=np.int64)
        s = np.where(a < L // 2, (a * 37) % (1 + a // 2) * np.where(a % 2, 1, -1), 0)
        s[L // 2 + 1::97] = 3
        return s
    if kind == "long_runs":
        s = np.zeros(L, np.int64)
        s[::1500] = rng.integers(1, 4, len(s[::1500]))
        s[-1] = 7
        return s
    if kind == "alternate":
        s = np.zeros(L, np.int64)
        s[::3] = rng.integers(-2, 3, len(s[::3]))
        return s
    if kind == "small":
        return rng.integers(-3, 4, L)
    if kind == "short_runs_tail":
        # runs of 1-3 zeros between small values keep k_P in the k = 1 and
        # k = 2 run states, so runs reach chunk ends there; the slice ends in
        # zeros, so the last run reaches the end of the slice
        s = rng.integers(1, 3, L) * rng.choice([-1, 1], L)
        s[rng.random(L) < 0.55] = 0
        s[-int(rng.integers(1, 4)):] = 0
        return s
    if kind == "k0_escape":
        # k = 0 mode (non-zero symbols keep k_P low) with sudden large values:
        # unary prefixes of 32+ (Golomb escapes) and codes over 32 bits
        s = rng.integers(1, 4, L) * rng.choice([-1, 1], L)
        m = rng.random(L) < 0.02
        s[m] = rng.integers(1 << 12, 1 << 30, m.sum()) * rng.choice([-1, 1], m.sum())
        return s
    raise ValueError(kind)


KINDS = ["sparse", "bursty", "huge", "mixed", "ramp", "long_runs", "alternate", "small", "k0_escape",
         "short_runs_tail"]


def _stream(pvvc, oracle, seed=12, per_kind=3):
    rng = np.random.default_rng(seed)
    slices, syms = [], []
    idx = 0
    for kind in KINDS:
        for _ in range(per_kind):
            n_target = int(rng.integers(300, 3000))
            v = np.unique(rng.integers(0, 64, (n_target, 3)), axis=0).astype(np.float32)
            geo, leaf, _ = pvvc.encode_colors_cpu(v, np.zeros((len(v), 3), np.float32), 1)
            n = len(leaf)
            s = _symbols(kind, 3 * n, rng)
            col = oracle.rlgr_encode(s)
            slices.append(dict(index=idx, type=0x01, size=n, qp=1, mv=IDENT, geo=geo, col=col))
            syms.append((kind, s))
            idx += 1
    return pack_frame(0, slices), syms


def test_stress_streams_round_trip_in_oracle(pvvc, oracle):
    """The crafted streams decode back to their symbols (CPU)."""
    img, syms = _stream(pvvc, oracle)
    f = pvvc.Frame.parse(img)
    for s, (kind, sym) in enumerate(syms):
        out, st = oracle.rlgr_decode(f.color(s), len(sym))
        assert st == 0, kind
        np.testing.assert_array_equal(np.asarray(out), sym, err_msg=kind)


@pytest.mark.gpu
def test_gpu_rlgr_stress_matches_oracle(pvvc, oracle, tmp_path):
    img, syms = _stream(pvvc, oracle)
    path = tmp_path / "stress_0.frame"
    path.write_bytes(img)
    ref = oracle.StreamDecoder(fp64=False).decode(oracle.OracleFrame(str(path)), 8)
    assert (ref["status"] == 0).all()
    f = pvvc.Frame.load(str(path))
    n_patches = f.slice_cnt
    max_pts = int(f.meta()["size"].max())
    want = ref["rgbf"] / np.float32(255)
    for waves in (2, 1):
        _say("rlgr stress drop-in, waves", waves)
        assert pvvc.hip_lib().pvvc_set_dropin_tuning(waves, 32 * 1024) == 0
        st = pvvc.State(n_patches, max_pts)
        pts = pvvc.decode_dropin(f, st)
        for s in range(n_patches):
            assert st.status(s) == 0, syms[s][0]
        np.testing.assert_array_equal(np.stack([pts["r"], pts["g"], pts["b"]], 1), want)
    assert pvvc.hip_lib().pvvc_set_dropin_tuning(4, 32 * 1024) == 0  # the default
    _say("rlgr stress persistent")
    dec = pvvc.Decoder(n_patches, max_pts)
    b = dec.batch([f])
    pts = dec.decode_host(b, pvvc.MODE_PERSISTENT, pvvc.OUT_POINTS)
    assert (b.status() == 0).all()
    np.testing.assert_array_equal(np.stack([pts["r"], pts["g"], pts["b"]], 1), want)
