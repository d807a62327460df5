"""Writes the golden fixtures used by tests/test_golden.py.

A two-frame stream (one intra frame, one predictive frame) of three hand-built
patches.  Its expected output comes from closed forms in FP64, not from any
decoder:

* patch 0 and patch 1 are octrees with header centre (1,1,1), range (2,2,2),
  height 2, and occupancy 0x81. Their two leaves sit at (1.5,1.5,1.5) and
  (0.5,0.5,0.5). The colour stream of each channel is [h, DC]. The leaves
  are (DC - h)/sqrt(2) and (DC + h)/sqrt(2).
* patch 2 has height 1: a single leaf at the header centre, coloured DC.

Motion is the identity for patch 0 and an integer translation for the other
two, so geometry stays exact.  The predictive frame adds its decoded
residual to the intra colours.  RGB follows BT.601 with the decoder's
constants, rounded.  Values are chosen so that no channel lies within 0.05 of
a rounding boundary, so FP32 decoding rounds the same way.  The intra blobs
of patch 2 are zstd frames (type bits 3 and 4).

Run from the repository root: python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes as C
import json
import math
import struct
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
sys.path.insert(0, str(REPO / "patchvvc-plus_amd"))
sys.path.insert(0, str(REPO / "tests"))

import pvvc  # noqa: E402
from _frames import pack_frame  # noqa: E402


def rlgr(sym) -> bytes:
    L = pvvc.enc_lib()
    v = np.asarray(sym, np.int64)
    n = L.pvvc_enc_rlgr(v.ctypes.data, v.size, None, 0)
    out = np.zeros(n, np.uint8)
    L.pvvc_enc_rlgr(v.ctypes.data, v.size, out.ctypes.data, n)
    return out.tobytes()


def zstd(data: bytes) -> bytes:
    lib = C.CDLL("libzstd.so.1")
    lib.ZSTD_compressBound.restype = C.c_size_t
    lib.ZSTD_compress.restype = C.c_size_t
    lib.ZSTD_compress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int]
    cap = lib.ZSTD_compressBound(C.c_size_t(len(data)))
    buf = C.create_string_buffer(cap)
    n = lib.ZSTD_compress(buf, cap, data, len(data), 3)
    return buf.raw[:n]


def header(centre, rng, height) -> bytes:
    return struct.pack(">6f", *centre, *rng) + bytes([height])


def translation(t):
    return [1, 0, 0, t[0], 0, 1, 0, t[1], 0, 0, 1, t[2], 0, 0, 0, 1]


def rgb_of(y, u, v):
    return (y + 1.4020 * (v - 128.0), y - 0.3441 * (u - 128.0) - 0.7141 * (v - 128.0), y + 1.7720 * (u - 128.0))


def main() -> None:
    qp = 2
    s2 = math.sqrt(2.0)
    # symbols per channel (h, DC) for the two-leaf patches, DC for the single leaf
    intra = {0: {"Y": (6, 129), "U": (-4, 90), "V": (3, 92)},
             1: {"Y": (-9, 75), "U": (5, 88), "V": (-2, 95)},
             2: {"Y": (None, 101), "U": (None, 58), "V": (None, 67)}}
    resid = {0: {"Y": (2, 9), "U": (-1, -3), "V": (4, 1)},
             1: {"Y": (0, -5), "U": (3, 2), "V": (-3, 6)},
             2: {"Y": (None, -9), "U": (None, 4), "V": (None, -6)}}
    geo2 = header((1, 1, 1), (2, 2, 2), 2) + bytes([0x81])
    geo1 = header((4, 5, 6), (1, 1, 1), 1)
    motion = {0: translation((0, 0, 0)), 1: translation((10, -3, 7)), 2: translation((-2, 8, 1))}
    leaves = {0: [(1.5, 1.5, 1.5), (0.5, 0.5, 0.5)], 1: [(1.5, 1.5, 1.5), (0.5, 0.5, 0.5)], 2: [(4.0, 5.0, 6.0)]}

    def stream(spec, single):
        sym = []
        for ch in "YUV":
            h, dc = spec[ch]
            sym += [dc] if single else [h, dc]
        return rlgr(sym)

    def yuv(spec, single):
        out = []
        for ch in "YUV":
            h, dc = spec[ch]
            if single:
                out.append([dc * qp])
            else:
                out.append([(dc * qp - h * qp) / s2, (dc * qp + h * qp) / s2])
        return np.array(out).T  # [leaf][channel]

    frames, expected = [], []
    ref = {}
    for k, table in enumerate((intra, resid)):
        slices, exp_xyz, exp_rgb = [], [], []
        for p in (0, 1, 2):
            single = p == 2
            col = stream(table[p], single)
            c = yuv(table[p], single)
            if k == 0:
                ref[p] = c
                ty = 0x01
                geo = geo1 if single else geo2
            else:
                c = c + ref[p]
                ty = 0x03
                geo = b""
            if k == 0 and p == 2:
                geo, col, ty = zstd(geo), zstd(col), ty | 0x08 | 0x10
            slices.append(dict(index=p, type=ty, size=len(leaves[p]), qp=qp, mv=motion[p], geo=geo, col=col))
            t = motion[p]
            for (x, y, z), (cy, cu, cv) in zip(leaves[p], c):
                exp_xyz.append([x + t[3], y + t[7], z + t[11]])
                rgb = rgb_of(cy, cu, cv)
                for v in rgb:
                    frac = v - math.floor(v)
                    assert abs(frac - 0.5) > 0.05 and 0.0 <= v <= 255.0, (k, p, rgb)
                exp_rgb.append([int(math.floor(v + 0.5)) for v in rgb])
        img = pack_frame(100 + k, slices)
        name = f"golden_{k}.frame"
        (HERE / name).write_bytes(img)
        frames.append(name)
        expected.append({"xyz": exp_xyz, "rgb": exp_rgb})
    (HERE / "golden_expected.json").write_text(json.dumps({"frames": frames, "expected": expected}, indent=1) + "\n")
    print("wrote", frames, "golden_expected.json")


if __name__ == "__main__":
    main()
