"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, never by the product.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
vp = C.c_void_p

_lib = None


def build() -> Path:
    r = subprocess.run(["make", "-C", str(HERE), "-s"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        build()
    L = C.CDLL(str(LIB))

    def sig(name, res, *args):
        f = getattr(L, name)
        f.restype = res
        f.argtypes = list(args)

    sig("orc_rlgr_encode", C.c_int64, vp, C.c_int64, vp, C.c_int64)
    sig("orc_rlgr_decode", C.c_int, vp, C.c_int64, C.c_int64, vp)
    sig("orc_patch_new", vp)
    sig("orc_patch_free", None, vp)
    sig("orc_patch_set_slice", C.c_int, vp, C.c_uint8, C.c_uint32, C.c_uint8, vp, C.c_uint32, vp, C.c_uint32, C.c_int)
    sig("orc_patch_get", C.c_int64, vp, vp, vp, vp, vp, vp)
    sig("orc_tree_expand", C.c_int64, vp, C.c_uint32, C.c_uint32, vp, vp)
    sig("orc_frame_load", vp, C.c_char_p, C.POINTER(C.c_int))
    sig("orc_frame_free", None, vp)
    sig("orc_frame_slice_cnt", C.c_uint32, vp)
    sig("orc_frame_timestamp", C.c_int, vp)
    sig("orc_frame_meta", None, vp, vp, vp, vp, vp, vp, vp, vp)
    sig("orc_frame_geo", vp, vp, C.c_uint32)
    sig("orc_frame_col", vp, vp, C.c_uint32)
    sig("orc_ctx_new", vp)
    sig("orc_ctx_free", None, vp)
    sig("orc_ctx_set_fp64", None, vp, C.c_int)
    sig("orc_decode_frame", C.c_int64, vp, vp, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp)
    sig("orc_cloud_mse", C.c_int, vp, vp, C.c_int64, vp, vp, C.c_int64, C.c_int, C.c_int, vp, vp, vp)
    sig("orc_raht_forward", C.c_int, vp, C.c_uint32, vp, C.c_uint32, vp)
    sig("orc_raht_inverse", C.c_int, vp, C.c_uint32, C.c_uint32, vp, vp)
    sig("orc_color_encode", C.c_int64, vp, C.c_uint32, vp, C.c_uint32, C.c_uint8, vp, vp, C.c_int64)
    sig("orc_save_slice", C.c_int64, C.c_char_p, C.c_uint8, C.c_int, C.c_int, C.c_uint64, C.c_uint8, vp, vp,
        C.c_uint64, vp, C.c_uint64)
    sig("orc_slices_to_frame", C.c_int, C.POINTER(C.c_char_p), C.c_int, C.c_char_p)
    _lib = L
    return L


def _buf(b: bytes):
    return (C.c_uint8 * max(1, len(b))).from_buffer_copy(b if b else b"\0")


def save_slice(path: str, type_: int, timestamp: int, index: int, size: int, qp: int, mv, geo: bytes | None,
               col: bytes | None) -> int:
    """SaveSlice (slice_io.cc:20-120); mv is 4x4 row-major.  Bytes written or minus a code."""
    m = np.ascontiguousarray(np.asarray(mv, np.float32).reshape(16))
    g = _buf(geo or b"") if geo else None
    c = _buf(col or b"") if col else None
    return lib().orc_save_slice(str(path).encode(), type_, timestamp, index, size, qp, m.ctypes.data, g,
                                len(geo or b""), c, len(col or b""))


def slices_to_frame(paths, out: str) -> int:
    """LoadSlice of each path, then ChangeSliceToFrame's .frame file."""
    arr = (C.c_char_p * max(1, len(paths)))(*[str(p).encode() for p in paths])
    return lib().orc_slices_to_frame(arr, len(paths), str(out).encode())


def rlgr_encode(sym) -> bytes:
    a = np.ascontiguousarray(sym, dtype=np.int64)
    n = lib().orc_rlgr_encode(a.ctypes.data, a.size, None, 0)
    out = np.zeros(max(1, n), np.uint8)
    lib().orc_rlgr_encode(a.ctypes.data, a.size, out.ctypes.data, n)
    return out[:n].tobytes()


def rlgr_decode(data: bytes, nsym: int) -> tuple[np.ndarray, int]:
    out = np.zeros(max(1, nsym), np.int64)
    st = lib().orc_rlgr_decode(_buf(data), len(data), nsym, out.ctypes.data)
    return out[:nsym], st


def raht_forward(geo: bytes, leaf_yuv) -> tuple[np.ndarray | None, int]:
    """Forward RAHT (raht_octree.cc:189-247) of leaf colours [n, 3] on intra
    geometry bytes: (coefficients [3, n] planar in stream order, status)."""
    y = np.ascontiguousarray(leaf_yuv, np.float32).reshape(-1, 3)
    n = len(y)
    c = np.zeros(3 * max(1, n), np.float32)
    st = lib().orc_raht_forward(_buf(geo), len(geo), y.ctypes.data, n, c.ctypes.data)
    return (c[:3 * n].reshape(3, n) if st == 0 else None), int(st)


def raht_inverse(geo: bytes, coef) -> tuple[np.ndarray | None, int]:
    """Inverse RAHT of planar coefficients [3, n]: (leaf colours [n, 3], status)."""
    c = np.ascontiguousarray(coef, np.float32).reshape(3, -1)
    n = c.shape[1]
    out = np.zeros((max(1, n), 3), np.float32)
    st = lib().orc_raht_inverse(_buf(geo), len(geo), n, c.ctypes.data, out.ctypes.data)
    return (out[:n] if st == 0 else None), int(st)


def color_encode(geo: bytes, leaf_yuv, qp: int) -> tuple[bytes | None, np.ndarray | None, int]:
    """Colour stream of one slice (GoP_encoding.cc:172-192, before zstd):
    (RLGR bytes, quantised symbols [3n] planar, status)."""
    y = np.ascontiguousarray(leaf_yuv, np.float32).reshape(-1, 3)
    n = len(y)
    q = np.zeros(3 * max(1, n), np.int64)
    r = lib().orc_color_encode(_buf(geo), len(geo), y.ctypes.data, n, qp, q.ctypes.data, None, 0)
    if r < 0:
        return None, None, int(-r)
    out = np.zeros(max(1, r), np.uint8)
    lib().orc_color_encode(_buf(geo), len(geo), y.ctypes.data, n, qp, None, out.ctypes.data, r)
    return out[:r].tobytes(), q[:3 * n], 0


def tree_expand(geo: bytes, size: int) -> tuple[np.ndarray | None, int, int]:
    """Leaf centres of intra geometry bytes: (xyz, status, height)."""
    xyz = np.zeros((max(1, size), 3), np.float32)
    h = C.c_int()
    r = lib().orc_tree_expand(_buf(geo), len(geo), size, xyz.ctypes.data, C.byref(h))
    if r < 0:
        return None, int(-r), h.value
    return xyz[:r], 0, h.value


class Patch:
    """One persistent per-patch decoder (InvertRAHTOctree)."""

    def __init__(self):
        self.h = lib().orc_patch_new()

    def __del__(self):
        try:
            lib().orc_patch_free(self.h)
        except Exception:
            pass

    def set_slice(self, type_: int, size: int, qp: int, geo: bytes, col: bytes, unwrap_zstd: bool = False) -> int:
        return lib().orc_patch_set_slice(self.h, type_, size, qp, _buf(geo) if geo else None, len(geo),
                                         _buf(col) if col else None, len(col), int(unwrap_zstd))

    def get(self, n: int) -> dict | None:
        xyz = np.zeros((n, 3), np.float32)
        yuv = np.zeros((n, 3), np.float32)
        rgb = np.zeros((n, 3), np.uint8)
        rgbf = np.zeros((n, 3), np.float32)
        rgb64 = np.zeros((n, 3), np.float64)
        r = lib().orc_patch_get(self.h, xyz.ctypes.data, yuv.ctypes.data, rgb.ctypes.data, rgbf.ctypes.data,
                                rgb64.ctypes.data)
        if r < 0:
            return None
        return dict(xyz=xyz[:r], yuv=yuv[:r], rgb=rgb[:r], rgbf=rgbf[:r], rgb64=rgb64[:r])


class OracleFrame:
    def __init__(self, path: str):
        rc = C.c_int()
        self.h = lib().orc_frame_load(str(path).encode(), C.byref(rc))
        self.rc = rc.value

    def __del__(self):
        try:
            if self.h:
                lib().orc_frame_free(self.h)
        except Exception:
            pass

    def meta(self) -> dict:
        S = lib().orc_frame_slice_cnt(self.h)
        m = {"index": np.zeros(S, np.int32), "type": np.zeros(S, np.uint8), "size": np.zeros(S, np.uint32),
             "qp": np.zeros(S, np.uint8), "mv": np.zeros((S, 16), np.float32),
             "geometry_size": np.zeros(S, np.uint32), "color_size": np.zeros(S, np.uint32)}
        lib().orc_frame_meta(self.h, *(m[k].ctypes.data for k in
                                       ("index", "type", "size", "qp", "mv", "geometry_size", "color_size")))
        m["timestamp"] = lib().orc_frame_timestamp(self.h)
        return m

    def geometry(self, s: int) -> bytes:
        m = self.meta()
        p = lib().orc_frame_geo(self.h, s)
        return C.string_at(p, int(m["geometry_size"][s])) if p else b""

    def color(self, s: int) -> bytes:
        m = self.meta()
        p = lib().orc_frame_col(self.h, s)
        return C.string_at(p, int(m["color_size"][s])) if p else b""


class StreamDecoder:
    """PVVCDecompression: persistent handlers per patch index, threaded."""

    def __init__(self, fp64: bool = True):
        self.h = lib().orc_ctx_new()
        lib().orc_ctx_set_fp64(self.h, int(fp64))

    def __del__(self):
        try:
            lib().orc_ctx_free(self.h)
        except Exception:
            pass

    def decode(self, frame: OracleFrame, threads: int = 8) -> dict:
        m = frame.meta()
        n = int(m["size"].astype(np.int64).sum())
        S = len(m["size"])
        out = dict(xyz=np.zeros((n, 3), np.float32), rgb=np.zeros((n, 3), np.uint8),
                   rgbf=np.zeros((n, 3), np.float32), rgb64=np.zeros((n, 3), np.float64),
                   leaf=np.zeros((n, 3), np.float32), status=np.zeros(max(1, S), np.int32))
        sat = C.c_int64()
        ms = np.zeros(2, np.float64)
        lib().orc_decode_frame(self.h, frame.h, threads, out["xyz"].ctypes.data, out["rgb"].ctypes.data,
                               out["rgbf"].ctypes.data, out["rgb64"].ctypes.data, out["leaf"].ctypes.data,
                               out["status"].ctypes.data, C.byref(sat), ms.ctypes.data)
        out["status"] = out["status"][:S]
        out["saturated"] = sat.value
        out["ms_setslice_getpatch"] = float(ms[0])
        out["ms_with_motion"] = float(ms[1])
        out["meta"] = m
        return out

    def decode_timing_only(self, frame: OracleFrame, threads: int, buffers: dict) -> tuple[float, float]:
        """Decode into preallocated buffers; returns (ms reference scope, ms incl. motion)."""
        sat = C.c_int64()
        ms = np.zeros(2, np.float64)
        lib().orc_decode_frame(self.h, frame.h, threads, buffers["xyz"].ctypes.data, buffers["rgb"].ctypes.data,
                               None, None, None, buffers["status"].ctypes.data, C.byref(sat), ms.ctypes.data)
        return float(ms[0]), float(ms[1])


MSE_BRUTE = 1
MSE_FLOAT_RUNNING = 2


def cloud_mse(p_xyz, p_rgb, q_xyz, q_rgb, mode: int = 0, threads: int = 8, want_nn: bool = False):
    """MSE::Compute restated (registration_base.cc:168-223).  Returns a dict
    geo/y/u/v -> (p->q, q->p) and, with want_nn, the neighbour indices."""
    px = np.ascontiguousarray(p_xyz, np.float32).reshape(-1, 3)
    qx = np.ascontiguousarray(q_xyz, np.float32).reshape(-1, 3)
    pc = np.ascontiguousarray(p_rgb, np.uint8).reshape(-1, 3)
    qc = np.ascontiguousarray(q_rgb, np.uint8).reshape(-1, 3)
    assert len(px) == len(pc) and len(qx) == len(qc)
    out = np.zeros(8, np.float64)
    nn_p = np.zeros(max(1, len(px)), np.int64)
    nn_q = np.zeros(max(1, len(qx)), np.int64)
    rc = lib().orc_cloud_mse(px.ctypes.data, pc.ctypes.data, len(px), qx.ctypes.data, qc.ctypes.data, len(qx),
                             mode, threads, out.ctypes.data, nn_p.ctypes.data, nn_q.ctypes.data)
    assert rc == 0
    res = {k: (float(out[i]), float(out[4 + i])) for i, k in enumerate(("geo", "y", "u", "v"))}
    if want_nn:
        res["nn_p"] = nn_p[: len(px)]
        res["nn_q"] = nn_q[: len(qx)]
    return res
