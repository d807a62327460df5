"""The lane RLGR decoder (rlgr_decode_lane, patchvvc-plus_amd/csrc/pvvc_lane.h),
built for the host and checked against the oracle's RLGR decoder
(entropy_codec.cc:207-283) on the CPU.

* Streams from the oracle's encoder over the crafted symbol kinds of the GPU
  stress test (runs with large exponents, escapes, k = 0 / run mode changes),
  whole and truncated, at every byte alignment of the colour bytes, with
  bytes after the end that must not be read: coefficients equal
  float(int64(symbol) * qp), statuses agree.
* Symbols up to the 32-bit range (Golomb escapes).
* Both dword sources: straight from memory, and through the LDS ring's
  indexing (rounds of 16 dwords, wrap-around, the spare row).
* Random bytes: every status class.  The GPU decoders reject a stored value
  outside 32 bits (SYMBOL_RANGE) where the oracle's 64-bit decoder carries
  on; everything else agrees with the oracle exactly.
"""
import ctypes as C
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from test_gpu_rlgr_stress import KINDS, _symbols

OK, BITSTREAM, SYMBOL_RANGE = 0, 5, 6
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def lane(tmp_path_factory):
    if not Path(HIPCC).exists():
        pytest.skip("hipcc not found")
    src = Path(__file__).with_name("lane_host.cpp")
    so = tmp_path_factory.mktemp("lane") / "lane_host.so"
    subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17", "-shared", "-fPIC",
                    str(src), "-o", str(so)], check=True, capture_output=True)
    lib = C.CDLL(str(so))
    for fn in (lib.lane_decode, lib.lane_decode_ring):
        fn.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]
        fn.restype = C.c_int

    def decode(data: bytes, total: int, qp: int, align: int):
        buf = np.full(len(data) + 12, 0xFF, np.uint8)  # bytes past the end read as ones would break runs
        buf[align:align + len(data)] = np.frombuffer(data, np.uint8)
        res = []
        for fn in (lib.lane_decode, lib.lane_decode_ring):
            out = np.zeros(max(total, 1), np.float32)
            st = fn(buf.ctypes.data + align, len(data), total, qp, out.ctypes.data)
            res.append((out[:total], st))
        (o1, s1), (o2, s2) = res
        assert s1 == s2 and (s1 != 0 or np.array_equal(o1, o2)), "memory and ring readers differ"
        return o1, s1
    return decode


def _mapped(s):
    s = np.asarray(s, np.int64)
    return np.where(s > 0, 2 * s.astype(np.float64) - 1, -2 * s.astype(np.float64))


def test_lane_decoder_matches_oracle_on_crafted_streams(lane, oracle):
    rng = np.random.default_rng(3)
    seen = set()
    for it in range(160):
        kind = KINDS[it % len(KINDS)]
        n = int(rng.integers(1, 5000))
        sym = _symbols(kind, n, rng)
        data = oracle.rlgr_encode(sym)
        qp = int(rng.integers(1, 40))
        for align in range(4):
            for cut in (len(data), len(data) // 2, len(data) - 1, 0):
                d = data[:cut]
                ref, rst = oracle.rlgr_decode(d, n)
                got, st = lane(d, n, qp, align)
                seen.add((rst, st))
                assert (st == OK) == (rst == OK), (kind, n, align, cut, st, rst)
                if st == OK:
                    np.testing.assert_array_equal(got, (ref * qp).astype(np.float32), err_msg=f"{kind} {n} {align}")
                else:
                    assert st == BITSTREAM == rst
    assert (OK, OK) in seen and (BITSTREAM, BITSTREAM) in seen


def test_lane_decoder_escapes_up_to_32_bits(lane, oracle):
    rng = np.random.default_rng(4)
    for it in range(120):
        n = int(rng.integers(1, 600))
        s = np.zeros(n, np.int64)
        m = rng.random(n) < rng.random()
        hi = [1 << 16, 1 << 30, (1 << 31) - 1, 1 << 31][it % 4]
        s[m] = rng.integers(-hi + 1, hi + 1, m.sum())
        data = oracle.rlgr_encode(s)
        got, st = lane(data, n, 1, it % 4)
        assert st == OK
        np.testing.assert_array_equal(got, s.astype(np.float32))


def test_lane_decoder_random_bytes(lane, oracle):
    rng = np.random.default_rng(9)
    classes = set()
    for it in range(6000):
        nbytes = int(rng.integers(0, 48))
        data = np.packbits(rng.random(8 * nbytes) < rng.random()).tobytes()
        n = int(rng.integers(1, 300))
        ref, rst = oracle.rlgr_decode(data, n)
        got, st = lane(data, n, 1, it % 4)
        out_of_range = bool((_mapped(ref) >= 2.0 ** 32).any())
        classes.add(st)
        if st == OK:
            assert rst == OK and not out_of_range
            np.testing.assert_array_equal(got, ref.astype(np.float32))
        elif st == BITSTREAM:
            assert rst == BITSTREAM
        else:
            assert st == SYMBOL_RANGE and (rst != OK or out_of_range)
        if rst == OK and not out_of_range:
            assert st == OK
    assert classes == {OK, BITSTREAM, SYMBOL_RANGE}
