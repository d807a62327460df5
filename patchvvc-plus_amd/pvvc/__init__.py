"""pvvc -- MI355X-native patchVVC per-patch GoP decoder (Python host layer).

The decoder itself is HIP (libpvvc_hip.so) behind the C ABI of
include/pvvc_hip.h; this module wraps that ABI with ctypes and numpy:

* Frame             -- a loaded .frame (io::LoadFrame, src/cuda/common/slice_io.cu:143-275)
* State, decode_dropin -- the reference's launch_cudaProcess call surface
                        (include/cuda/render.cuh:41-57) with its device arrays
* Decoder, Batch    -- native batched decode (per-frame, persistent, graph)
* synth_stream      -- seeded synthetic humanoid streams (stream writer)
* quality           -- MSE / PSNR of decoded clouds (vvc::common::MSE)
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Iterable, Sequence

import numpy as np

from . import _native
from ._native import DropinFrame, Points, SliceDesc, SynthConfig, Tuning, enc_lib, hip_lib

__all__ = [
    "PvvcError", "StagedFrame", "launch_dropin", "synchronize", "set_timing", "last_kernel_ms", "last_launch_times",
    "STATUS", "Frame", "DeviceBuffer", "State", "Decoder", "Batch", "decode_dropin", "synth_stream",
    "SYNTH_PRESETS", "Points", "hip_lib", "enc_lib", "MODE_PER_FRAME", "MODE_PERSISTENT", "MODE_GRAPH",
    "OUT_NATIVE", "OUT_POINTS", "device_info", "encode_colors_packed", "Stream",
]

STATUS = {
    0: "OK", 1: "BAD_SLICE", 2: "EMPTY_REFERENCE", 3: "UNMATCHED_SIZE", 4: "ZSTD", 5: "BITSTREAM",
    6: "SYMBOL_RANGE", 7: "CAPACITY", 8: "INVALID_ARG", 9: "HIP", 10: "TIMEOUT", 11: "IO", 12: "BUSY",
}
MODE_PER_FRAME, MODE_PERSISTENT, MODE_GRAPH = 0, 1, 2
OUT_NATIVE, OUT_POINTS = 0, 1
POINTS_DTYPE = np.dtype([(n, "<f4") for n in ("x", "y", "z", "r", "g", "b")])


class PvvcError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        super().__init__(f"{what}: {STATUS.get(code, code)} ({code})")
        self.code = code


def _check(code: int, what: str) -> None:
    if code != 0:
        raise PvvcError(code, what)


def device_info() -> dict:
    L = hip_lib()
    cu, lds, warp = C.c_int(), C.c_int(), C.c_int()
    name = C.create_string_buffer(256)
    _check(L.pvvc_device_info(C.byref(cu), C.byref(lds), C.byref(warp), name, 256), "device_info")
    return {"cu": cu.value, "lds_per_block": lds.value, "warp": warp.value, "arch": name.value.decode()}


# ---------------------------------------------------------------------------
# Frames
# ---------------------------------------------------------------------------
class Frame:
    """A .frame loaded and zstd-unwrapped on the host."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    @classmethod
    def load(cls, path: str) -> "Frame":
        h = C.c_void_p()
        rc = hip_lib().pvvc_load_frame(str(path).encode(), C.byref(h))
        if rc != 0:
            raise PvvcError(rc, f"pvvc_load_frame({path})")
        return cls(h.value)

    @classmethod
    def parse(cls, data: bytes) -> "Frame":
        h = C.c_void_p()
        buf = (C.c_uint8 * len(data)).from_buffer_copy(data) if data else None
        rc = hip_lib().pvvc_parse_frame(buf, len(data), C.byref(h))
        if rc != 0:
            raise PvvcError(rc, "pvvc_parse_frame")
        return cls(h.value)

    @classmethod
    def from_slices(cls, paths) -> "Frame":
        """The frame of .slice files (pvvc_load_slices: LoadSlice + ChangeSliceToFrame)."""
        h = C.c_void_p()
        arr = (C.c_char_p * max(1, len(paths)))(*[str(p).encode() for p in paths])
        rc = hip_lib().pvvc_load_slices(arr, len(paths), C.byref(h))
        if rc != 0:
            raise PvvcError(rc, "pvvc_load_slices")
        return cls(h.value)

    def close(self) -> None:
        if self._h and self._h.value:
            hip_lib().pvvc_frame_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @property
    def slice_cnt(self) -> int:
        return int(hip_lib().pvvc_frame_slice_cnt(self._h))

    @property
    def points(self) -> int:
        return int(hip_lib().pvvc_frame_points(self._h))

    @property
    def timestamp(self) -> int:
        return int(hip_lib().pvvc_frame_timestamp(self._h))

    def meta(self) -> dict:
        S = self.slice_cnt
        out = {
            "index": np.zeros(S, np.int32), "type": np.zeros(S, np.uint8), "size": np.zeros(S, np.uint32),
            "qp": np.zeros(S, np.uint8), "mv": np.zeros((S, 16), np.float32),
            "geometry_size": np.zeros(S, np.uint32), "color_size": np.zeros(S, np.uint32),
        }
        hip_lib().pvvc_frame_meta(self._h, *(out[k].ctypes.data for k in
                                             ("index", "type", "size", "qp", "mv", "geometry_size", "color_size")))
        return out

    def geometry(self, s: int) -> bytes:
        m = self.meta()
        p = hip_lib().pvvc_frame_geometry(self._h, s)
        return C.string_at(p, int(m["geometry_size"][s])) if p else b""

    def color(self, s: int) -> bytes:
        m = self.meta()
        p = hip_lib().pvvc_frame_color(self._h, s)
        return C.string_at(p, int(m["color_size"][s])) if p else b""


# ---------------------------------------------------------------------------
# Device memory
# ---------------------------------------------------------------------------

SLICE_CODES = {1: "wrong file name", 2: "no such file", 3: "permission denied", 4: "open error / not a directory",
               5: "short read", 6: "zstd", 7: "slice type without the valid bit", 8: "empty blob", 9: "no slices",
               10: "write error"}


def save_slice(path, type_: int, timestamp: int, index: int, size: int, qp: int, mv, geometry: bytes | None = None,
               color: bytes | None = None) -> int:
    """pvvc_save_slice (io::SaveSlice): one .slice file; mv is 4x4 row-major.  Returns the bytes written."""
    d = SliceDesc()
    d.type, d.timestamp, d.index, d.size, d.qp = type_, timestamp, index, size, qp
    d.mv[:] = [float(x) for x in np.asarray(mv, np.float32).reshape(16)]
    keep = []
    for name, blob in (("geometry", geometry), ("color", color)):
        if blob:
            b = (C.c_uint8 * len(blob)).from_buffer_copy(blob)
            keep.append(b)
            setattr(d, name, C.cast(b, C.c_void_p).value)
            setattr(d, name + "_size", len(blob))
    n = hip_lib().pvvc_save_slice(str(path).encode(), C.byref(d))
    if n < 0:
        raise PvvcError(-n, f"pvvc_save_slice({path}): {SLICE_CODES.get(-n, '')}")
    return int(n)


def slices_to_frame_file(paths, frame_path) -> None:
    """pvvc_slices_to_frame_file: the .frame file ChangeSliceToFrame writes for these slices."""
    arr = (C.c_char_p * max(1, len(paths)))(*[str(p).encode() for p in paths])
    _check(hip_lib().pvvc_slices_to_frame_file(arr, len(paths), str(frame_path).encode()), "pvvc_slices_to_frame_file")


def change_slice_to_frame(input_dir, frame_path) -> None:
    """pvvc_change_slice_to_frame (io::ChangeSliceToFrame over a directory)."""
    _check(hip_lib().pvvc_change_slice_to_frame(str(input_dir).encode(), str(frame_path).encode()),
           "pvvc_change_slice_to_frame")


class DeviceBuffer:
    """A plain hipMalloc allocation."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = hip_lib().pvvc_dev_alloc(max(1, self.nbytes))
        if not self.ptr:
            raise PvvcError(9, f"pvvc_dev_alloc({nbytes})")

    def free(self) -> None:
        if self.ptr:
            hip_lib().pvvc_dev_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def zero(self) -> None:
        _check(hip_lib().pvvc_memset(C.c_void_p(self.ptr), 0, self.nbytes), "memset")

    def read(self, dtype, count: int, offset_bytes: int = 0) -> np.ndarray:
        a = np.empty(count, dtype=dtype)
        if a.nbytes:
            _check(hip_lib().pvvc_memcpy_d2h(a.ctypes.data, C.c_void_p(self.ptr + offset_bytes), a.nbytes), "d2h")
        return a

    def write(self, arr: np.ndarray, offset_bytes: int = 0) -> None:
        a = np.ascontiguousarray(arr)
        _check(hip_lib().pvvc_memcpy_h2d(C.c_void_p(self.ptr + offset_bytes), a.ctypes.data, a.nbytes), "h2d")


def synchronize() -> None:
    _check(hip_lib().pvvc_synchronize(), "synchronize")


def set_timing(on: bool) -> bool:
    """Kernel timing events for this thread's decode calls (off by default);
    returns the previous setting."""
    return bool(hip_lib().pvvc_set_timing(1 if on else 0))


def last_kernel_ms() -> float:
    return float(hip_lib().pvvc_last_kernel_ms())


def set_host_threads(n: int) -> None:
    """Host threads used to unwrap a frame's zstd blobs."""
    _check(hip_lib().pvvc_set_host_threads(int(n)), "pvvc_set_host_threads")


def last_launch_times() -> list[float]:
    L = hip_lib()
    n = L.pvvc_last_launch_times(None, 0)
    a = np.zeros(max(n, 1), np.float64)
    L.pvvc_last_launch_times(a.ctypes.data, n)
    return a[:n].tolist()


def last_decode_kernel_times() -> list[float]:
    """Durations (ms) of the decode kernels alone in the last decode call."""
    L = hip_lib()
    n = L.pvvc_last_decode_kernel_times(None, 0)
    a = np.zeros(max(n, 1), np.float64)
    L.pvvc_last_decode_kernel_times(a.ctypes.data, n)
    return a[:n].tolist()


# ---------------------------------------------------------------------------
# Drop-in path
# ---------------------------------------------------------------------------
class State:
    """Device state slots replacing the reference's InvertRAHTOctree array."""

    def __init__(self, n_patches: int, max_points: int):
        self.n_patches = n_patches
        self.max_points = max_points
        self.ptr = hip_lib().pvvc_state_alloc(n_patches, max_points)
        if not self.ptr:
            raise PvvcError(hip_lib().pvvc_last_error() or 9, "pvvc_state_alloc")

    def free(self) -> None:
        if self.ptr:
            hip_lib().pvvc_state_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def status(self, patch: int) -> int:
        return int(hip_lib().pvvc_state_patch_status(C.c_void_p(self.ptr), patch))

    def set_stream(self, stream: "Stream | None") -> None:
        """This state's own stream for launch_cudaProcess (None: the null stream)."""
        _check(hip_lib().pvvc_state_set_stream(C.c_void_p(self.ptr), C.c_void_p(stream.handle if stream else None)),
               "pvvc_state_set_stream")

    def set_dropin_tuning(self, waves: int, lds_bytes: int) -> None:
        _check(hip_lib().pvvc_state_set_dropin_tuning(C.c_void_p(self.ptr), waves, lds_bytes),
               "pvvc_state_set_dropin_tuning")

    def synchronize(self) -> int:
        """Waits for the state's stream; returns the first failed slice's status since the last read."""
        return int(hip_lib().pvvc_state_synchronize(C.c_void_p(self.ptr)))


class Stream:
    """A non-blocking HIP stream at the default priority (pvvc_stream_create)."""

    def __init__(self):
        h = C.c_void_p()
        _check(hip_lib().pvvc_stream_create(C.byref(h)), "pvvc_stream_create")
        self.handle = h.value

    def synchronize(self) -> None:
        _check(hip_lib().pvvc_stream_synchronize(C.c_void_p(self.handle)), "pvvc_stream_synchronize")

    def close(self) -> None:
        if self.handle:
            hip_lib().pvvc_stream_destroy(C.c_void_p(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class StagedFrame:
    """Device arrays of one drop-in call (pvvc_dropin_stage)."""

    def __init__(self, frame: Frame):
        self.d = DropinFrame()
        _check(hip_lib().pvvc_dropin_stage(frame.handle, C.byref(self.d)), "pvvc_dropin_stage")

    def release(self) -> None:
        if self.d.blob:
            hip_lib().pvvc_dropin_release(C.byref(self.d))

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Stager:
    """Reusable drop-in staging buffers: one device buffer and one pinned
    host buffer (pvvc_dropin_stage_to); stage() allocates nothing."""

    class Staged:
        def __init__(self, d: DropinFrame):
            self.d = d

    def __init__(self, capacity: int):
        self.capacity = int(capacity)
        self.dev = DeviceBuffer(self.capacity)
        self.host = hip_lib().pvvc_host_alloc(self.capacity)
        if not self.host:
            raise PvvcError(9, "pvvc_host_alloc")

    @staticmethod
    def bytes_for(frame: Frame) -> int:
        return int(hip_lib().pvvc_dropin_stage_bytes(frame.handle))

    def stage(self, frame: Frame) -> "Stager.Staged":
        d = DropinFrame()
        _check(hip_lib().pvvc_dropin_stage_to(frame.handle, C.c_void_p(self.dev.ptr), self.capacity,
                                              C.c_void_p(self.host), self.capacity, C.byref(d)),
               "pvvc_dropin_stage_to")
        return Stager.Staged(d)

    def close(self) -> None:
        if self.host:
            hip_lib().pvvc_host_free(C.c_void_p(self.host))
            self.host = None
        self.dev.free()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def launch_dropin(staged: StagedFrame, state: State, out: DeviceBuffer, timestamp: int = 0) -> None:
    """launch_cudaProcess with the staged arrays (asynchronous)."""
    d = staged.d
    hip_lib().launch_cudaProcess(1000, 1, C.c_void_p(out.ptr), timestamp, d.inner_offset, d.index, d.type, d.mv,
                                 d.size, d.qp, d.geometry, d.geometry_size, d.color, d.color_size,
                                 C.c_void_p(state.ptr), d.slice_cnt)


def decode_dropin(frame: Frame, state: State, out: DeviceBuffer | None = None) -> np.ndarray:
    """Stages, decodes and reads back one frame as reference Points."""
    staged = StagedFrame(frame)
    n = int(staged.d.point_cnt)
    own = out is None
    if own:
        out = DeviceBuffer(max(1, n) * 24)
    launch_dropin(staged, state, out, frame.timestamp)
    state.synchronize()  # the state's stream; launch errors go to this thread's pvvc_last_error
    res = out.read(POINTS_DTYPE, n)
    staged.release()
    if own:
        out.free()
    return res


# ---------------------------------------------------------------------------
# Native decoder
# ---------------------------------------------------------------------------
class Batch:
    def __init__(self, dec: "Decoder", frames: Sequence[Frame]):
        self.dec = dec
        arr = (C.c_void_p * len(frames))(*[f.handle.value for f in frames])
        h = C.c_void_p()
        _check(hip_lib().pvvc_batch_create(dec.handle, arr, len(frames), C.byref(h)), "pvvc_batch_create")
        self.handle = h
        self.n_frames = len(frames)
        L = hip_lib()
        self.frame_points = [int(L.pvvc_batch_frame_points(h, f)) for f in range(self.n_frames)]
        self.frame_offset = [int(L.pvvc_batch_frame_offset(h, f)) for f in range(self.n_frames)]
        self.total_points = int(sum(self.frame_points))
        self.input_bytes = int(L.pvvc_batch_input_bytes(h))
        self.state_bytes = int(L.pvvc_batch_state_bytes(h))
        self.items = int(L.pvvc_batch_items(h))

    def free(self) -> None:
        if self.handle and self.handle.value:
            hip_lib().pvvc_batch_free(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def status(self) -> np.ndarray:
        s = np.zeros(max(1, self.items), np.int32)
        _check(hip_lib().pvvc_batch_status(self.handle, s.ctypes.data, self.items), "pvvc_batch_status")
        return s[: self.items]


@dataclass
class Outputs:
    xyz: DeviceBuffer | None = None
    rgb: DeviceBuffer | None = None
    pts: DeviceBuffer | None = None
    points: int = 0


class Decoder:
    def __init__(self, max_patches: int, max_points: int, tuning: dict | None = None):
        h = C.c_void_p()
        _check(hip_lib().pvvc_decoder_create(C.byref(h), max_patches, max_points), "pvvc_decoder_create")
        self.handle = h
        if tuning:
            self.set_tuning(**tuning)

    def close(self) -> None:
        if self.handle and self.handle.value:
            hip_lib().pvvc_decoder_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, waves_per_patch=0, lds_bytes=0, wgs_per_cu=0, queue_chunk=0, priority_items=0,
                   lane_max_points=0) -> None:
        """Zero fields keep their value; lane_max_points=-1 turns lane decoding off."""
        t = Tuning(waves_per_patch, lds_bytes, wgs_per_cu, queue_chunk, priority_items, lane_max_points)
        _check(hip_lib().pvvc_decoder_set_tuning(self.handle, C.byref(t)), "set_tuning")

    def set_stream(self, stream: "Stream | None") -> None:
        """The decoder's own stream for its decodes (None: the null stream)."""
        _check(hip_lib().pvvc_decoder_set_stream(self.handle, C.c_void_p(stream.handle if stream else None)),
               "pvvc_decoder_set_stream")

    def set_white_small_patches(self, enable: int) -> None:
        _check(hip_lib().pvvc_decoder_set_white_small_patches(self.handle, int(enable)),
               "pvvc_decoder_set_white_small_patches")

    def tuning(self) -> dict:
        t = Tuning()
        _check(hip_lib().pvvc_decoder_get_tuning(self.handle, C.byref(t)), "get_tuning")
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    def reset(self) -> None:
        _check(hip_lib().pvvc_decoder_reset(self.handle), "reset")

    def batch(self, frames: Sequence[Frame]) -> Batch:
        return Batch(self, frames)

    def outputs(self, batch: Batch, output: int = OUT_NATIVE) -> Outputs:
        n = max(1, batch.total_points)
        if output == OUT_NATIVE:
            return Outputs(xyz=DeviceBuffer(12 * n), rgb=DeviceBuffer(3 * n), points=batch.total_points)
        return Outputs(pts=DeviceBuffer(24 * n), points=batch.total_points)

    def decode(self, batch: Batch, mode: int, out: Outputs) -> None:
        """Asynchronous decode of the batch into `out`."""
        out_mode = OUT_NATIVE if out.xyz is not None else OUT_POINTS
        p = lambda b: C.c_void_p(b.ptr) if b is not None else None  # noqa: E731
        _check(hip_lib().pvvc_decode_batch(self.handle, batch.handle, mode, out_mode, p(out.xyz), p(out.rgb),
                                           p(out.pts)), "pvvc_decode_batch")

    def autotune(self, batch: Batch, mode: int = MODE_PERSISTENT, out: Outputs | None = None,
                 reps: int = 3) -> tuple[dict, float]:
        """Runs the native autotuner on `batch`; returns (tuning, ms per decode call)
        and leaves the decoder set to that tuning."""
        own = out is None
        if own:
            out = self.outputs(batch, OUT_NATIVE)
        out_mode = OUT_NATIVE if out.xyz is not None else OUT_POINTS
        p = lambda b: C.c_void_p(b.ptr) if b is not None else None  # noqa: E731
        t = Tuning()
        ms = C.c_double()
        _check(hip_lib().pvvc_decoder_autotune(self.handle, batch.handle, mode, out_mode, p(out.xyz), p(out.rgb),
                                               p(out.pts), reps, C.byref(t), C.byref(ms)), "pvvc_decoder_autotune")
        return {k: getattr(t, k) for k, _ in Tuning._fields_}, float(ms.value)

    def outputs_for(self, frames: Sequence[Frame], output: int = OUT_NATIVE) -> Outputs:
        n = sum(f.points for f in frames)
        m = max(1, n)
        if output == OUT_NATIVE:
            return Outputs(xyz=DeviceBuffer(12 * m), rgb=DeviceBuffer(3 * m), points=n)
        return Outputs(pts=DeviceBuffer(24 * m), points=n)

    def decode_frame(self, frame: Frame, out: Outputs) -> tuple[int, np.ndarray]:
        """pvvc_decode_frame: one frame, synchronous.  Returns (code, per-slice
        statuses); code is the first non-OK status or 0."""
        st = np.zeros(max(1, frame.slice_cnt), np.int32)
        p = lambda b: C.c_void_p(b.ptr) if b is not None else None  # noqa: E731
        rc = hip_lib().pvvc_decode_frame(self.handle, frame.handle, OUT_NATIVE if out.xyz is not None else OUT_POINTS,
                                         p(out.xyz), p(out.rgb), p(out.pts), st.ctypes.data, frame.slice_cnt)
        return int(rc), st[:frame.slice_cnt]

    def decode_gop(self, frames: Sequence[Frame], out: Outputs) -> tuple[int, np.ndarray]:
        """pvvc_decode_gop: frames in decode order, one persistent launch,
        synchronous.  Returns (code, per-slice statuses in frame order)."""
        n = sum(f.slice_cnt for f in frames)
        st = np.zeros(max(1, n), np.int32)
        arr = (C.c_void_p * len(frames))(*[f.handle.value for f in frames])
        p = lambda b: C.c_void_p(b.ptr) if b is not None else None  # noqa: E731
        rc = hip_lib().pvvc_decode_gop(self.handle, arr, len(frames), OUT_NATIVE if out.xyz is not None else OUT_POINTS,
                                       p(out.xyz), p(out.rgb), p(out.pts), st.ctypes.data, n)
        return int(rc), st[:n]

    def decode_host(self, batch: Batch, mode: int = MODE_PERSISTENT, output: int = OUT_NATIVE):
        """Decodes and reads back: (xyz f32[N,3], rgb u8[N,3]) or Points[N]."""
        out = self.outputs(batch, output)
        self.decode(batch, mode, out)
        synchronize()
        n = batch.total_points
        if output == OUT_NATIVE:
            return out.xyz.read(np.float32, 3 * n).reshape(n, 3), out.rgb.read(np.uint8, 3 * n).reshape(n, 3)
        return out.pts.read(POINTS_DTYPE, n)


# ---------------------------------------------------------------------------
# Forward path (colour encoder)
# ---------------------------------------------------------------------------
def _color_job_dtype() -> np.dtype:
    """numpy view of pvvc_color_job (the ctypes ColorJob layout)."""
    from ._native import ColorJob
    names = [f for f, _ in ColorJob._fields_]
    fmt = {"geometry": "<u8", "geometry_size": "<u4", "yuv": "<u8", "n": "<u4", "qp": "u1", "out": "<u8",
           "out_cap": "<u4", "out_size": "<u4", "status": "<i4"}
    return np.dtype({"names": names, "formats": [fmt[n] for n in names],
                     "offsets": [getattr(ColorJob, n).offset for n in names], "itemsize": C.sizeof(ColorJob)})


def encode_colors(jobs: Sequence[tuple[bytes, np.ndarray, int]]) -> tuple[list[bytes], list[int], float]:
    """GPU colour encoder: jobs of (octree bytes, leaf YUV float32 [n, 3], qp).
    Returns (RLGR byte strings, statuses, kernel ms).  The job table is
    filled in one vectorised pass over three concatenated buffers."""
    n = len(jobs)
    if n == 0:
        _check(hip_lib().pvvc_encode_colors(None, 0, None), "pvvc_encode_colors")
    geos = [bytes(j[0]) for j in jobs]
    yuvs = [np.ascontiguousarray(j[1], np.float32).reshape(-1, 3) for j in jobs]
    gsz = np.fromiter((len(g) for g in geos), np.int64, n)
    npts = np.fromiter((len(y) for y in yuvs), np.int64, n)
    qp = np.fromiter((int(j[2]) for j in jobs), np.uint8, n)
    out_all, ooff, sizes, status, ms = encode_colors_packed(
        np.frombuffer(b"".join(geos), np.uint8), gsz, np.concatenate(yuvs), npts, qp)
    outs = [out_all[o:o + z].tobytes() for o, z in zip(ooff.tolist(), sizes.tolist())]
    return outs, status.tolist(), ms


def encode_colors_packed(geo_all: np.ndarray, geo_size: np.ndarray, yuv_all: np.ndarray, npts: np.ndarray,
                         qp: np.ndarray) -> tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray, float]:
    """GPU colour encoder over packed inputs: job i's octree bytes are the next
    geo_size[i] bytes of geo_all, its leaf YUV the next npts[i] rows of
    yuv_all (float32 [N, 3]), its quantisation step qp[i].  Returns
    (out_all, out_off, out_size, status, kernel ms): job i's RLGR bytes are
    out_all[out_off[i]: out_off[i] + out_size[i]].  The job table is filled by
    array arithmetic alone, and the output buffer is not cleared (the library
    writes exactly out_size[i] bytes per job)."""
    geo_all = np.ascontiguousarray(geo_all, np.uint8)
    yuv_all = np.ascontiguousarray(yuv_all, np.float32).reshape(-1, 3)
    gsz = np.asarray(geo_size, np.int64)
    npts = np.asarray(npts, np.int64)
    n = len(npts)
    if n == 0 or len(gsz) != n or len(qp) != n:
        raise PvvcError(8, "encode_colors_packed: job arrays of different lengths (or none)")
    if int(gsz.sum()) > geo_all.size or int(npts.sum()) > len(yuv_all):
        raise PvvcError(8, "encode_colors_packed: sizes exceed the packed inputs")
    goff = np.zeros(n, np.int64)
    np.cumsum(gsz[:-1], out=goff[1:])
    poff = np.zeros(n, np.int64)
    np.cumsum(npts[:-1], out=poff[1:])
    cap = 24 * npts + 64
    ooff = np.zeros(n, np.int64)
    np.cumsum(cap[:-1], out=ooff[1:])
    out_all = np.empty(int(cap.sum()), np.uint8)
    table = np.zeros(n, _color_job_dtype())
    table["geometry"] = geo_all.ctypes.data + goff
    table["geometry_size"] = gsz
    table["yuv"] = yuv_all.ctypes.data + 12 * poff
    table["n"] = npts
    table["qp"] = qp
    table["out"] = out_all.ctypes.data + ooff
    table["out_cap"] = cap
    ms = C.c_double()
    _check(hip_lib().pvvc_encode_colors(C.cast(table.ctypes.data, C.c_void_p), n, C.byref(ms)),
           "pvvc_encode_colors")
    return out_all, ooff, table["out_size"].astype(np.int64), table["status"].copy(), float(ms.value)


def encode_colors_cpu_on_tree(geo: bytes, leaf_yuv: np.ndarray, qp: int) -> bytes:
    """The stream writer's CPU encoder on a given octree (bytes) and leaf YUV."""
    L = enc_lib()
    g = np.frombuffer(bytes(geo), np.uint8).copy()
    y = np.ascontiguousarray(leaf_yuv, np.float32).reshape(-1, 3)
    out = np.zeros(24 * len(y) + 64, np.uint8)
    r = L.pvvc_enc_colors_on_tree(g.ctypes.data, g.size, y.ctypes.data, len(y), int(qp), out.ctypes.data, out.size)
    if r < 0:
        raise PvvcError(1, "pvvc_enc_colors_on_tree")
    return out[:r].tobytes()


def encode_colors_cpu(xyz: np.ndarray, yuv: np.ndarray, qp: int) -> tuple[bytes, np.ndarray, bytes]:
    """The stream writer's CPU encoder of one patch: (octree bytes, leaf YUV, RLGR bytes)."""
    L = enc_lib()
    p = np.ascontiguousarray(xyz, np.float32)
    c = np.ascontiguousarray(yuv, np.float32)
    n = len(p)
    geo = np.zeros(64 * n + 128, np.uint8)  # internal nodes <= (height - 1) x n
    gsz = C.c_int64()
    leaf = np.zeros((n, 3), np.float32)
    nl = C.c_int()
    col = np.zeros(24 * n + 64, np.uint8)
    r = L.pvvc_enc_patch_colors(p.ctypes.data, n, c.ctypes.data, int(qp), geo.ctypes.data, geo.size, C.byref(gsz),
                                leaf.ctypes.data, C.byref(nl), col.ctypes.data, col.size)
    if r < 0:
        raise PvvcError(1, "pvvc_enc_patch_colors")
    return geo[: gsz.value].tobytes(), leaf[: nl.value], col[:r].tobytes()


# ---------------------------------------------------------------------------
# Serving loop
# ---------------------------------------------------------------------------
STREAM_WAITING, STREAM_ACTIVE, STREAM_CLOSING, STREAM_CLOSED = 0, 1, 2, 3


class Server:
    """Multi-stream decode server (continuous batching, HBM admission).

    Frames passed to submit() are kept referenced here until their
    completion has been polled."""

    def __init__(self, hbm_budget: int = 0, lanes: int = 0, max_slots: int = 0, max_points_per_patch: int = 0,
                 out_pool_points: int = 0, max_batch_points: int = 0, tuning: dict | None = None,
                 lanes_spin: bool = True):
        from ._native import ServerConfig
        t = Tuning(**(tuning or {}))
        cfg = ServerConfig(hbm_budget, lanes, max_slots, max_points_per_patch, out_pool_points, max_batch_points, t,
                           int(lanes_spin))
        h = C.c_void_p()
        _check(hip_lib().pvvc_server_create(C.byref(cfg), C.byref(h)), "pvvc_server_create")
        self.handle = h
        self._held: dict[tuple[int, int], Frame] = {}
        self._seq: dict[int, int] = {}

    @staticmethod
    def clock() -> float:
        return float(hip_lib().pvvc_server_clock())

    def close(self) -> None:
        if self.handle and self.handle.value:
            hip_lib().pvvc_server_destroy(self.handle)
            self.handle = C.c_void_p()
        self._held.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def open_stream(self, max_patches: int, max_points_per_patch: int, max_points_per_frame: int,
                    out_ring: int = 2) -> int:
        sid = C.c_int()
        _check(hip_lib().pvvc_server_open_stream(self.handle, max_patches, max_points_per_patch, max_points_per_frame,
                                                 out_ring, C.byref(sid)), "pvvc_server_open_stream")
        self._seq[sid.value] = 0
        return sid.value

    def stream_state(self, sid: int) -> int:
        return int(hip_lib().pvvc_server_stream_state(self.handle, sid))

    def submit(self, sid: int, frame: Frame, arrival: float | None = None, tag: int = 0) -> int:
        """Queues a frame; returns its sequence number within the stream."""
        if arrival is None:
            arrival = self.clock()
        _check(hip_lib().pvvc_server_submit(self.handle, sid, frame.handle, arrival, tag), "pvvc_server_submit")
        seq = self._seq[sid]
        self._seq[sid] = seq + 1
        self._held[(sid, seq)] = frame
        return seq

    def close_stream(self, sid: int) -> None:
        _check(hip_lib().pvvc_server_close_stream(self.handle, sid), "pvvc_server_close_stream")

    def poll(self, max_n: int = 256, timeout_ms: int = 100) -> list[dict]:
        from ._native import Completion
        buf = (Completion * max_n)()
        k = hip_lib().pvvc_server_poll(self.handle, buf, max_n, timeout_ms)
        if k < 0:
            raise PvvcError(-k, "pvvc_server_poll")
        out = []
        for c in buf[:k]:
            self._held.pop((c.stream_id, c.seq), None)
            out.append({f: getattr(c, f) for f, _ in Completion._fields_})
        return out

    def read(self, comp: dict) -> tuple[np.ndarray, np.ndarray]:
        """Copies a completion's output to the host: (xyz f32[N,3], rgb u8[N,3])."""
        n = int(comp["points"])
        xyz = np.empty((n, 3), np.float32)
        rgb = np.empty((n, 3), np.uint8)
        if n:
            _check(hip_lib().pvvc_memcpy_d2h(xyz.ctypes.data, C.c_void_p(comp["xyz"]), 12 * n), "d2h")
            _check(hip_lib().pvvc_memcpy_d2h(rgb.ctypes.data, C.c_void_p(comp["rgb"]), 3 * n), "d2h")
        return xyz, rgb

    def release(self, comp: dict) -> None:
        _check(hip_lib().pvvc_server_release(self.handle, comp["stream_id"], comp["out_slot"]), "release")

    def stats(self) -> dict:
        from ._native import ServerStats
        st = ServerStats()
        _check(hip_lib().pvvc_server_get_stats(self.handle, C.byref(st)), "pvvc_server_get_stats")
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def reset_stats(self) -> None:
        _check(hip_lib().pvvc_server_reset_stats(self.handle), "reset_stats")


# ---------------------------------------------------------------------------
# Synthetic streams (stream writer)
# ---------------------------------------------------------------------------
# patch_min / patch_max / patch_mu / patch_sigma drive the recursive median
# split: a group stops splitting at a lognormal target clamped to
# [patch_min, patch_max], so final patches fall between about half the
# target and the target.  Measured, first frame, seed 7 (5, 6 for c2/c4):
#   c0 121 856 points, 123 patches of 669-1993 (median 996)
#   c1 816 807 points, 866 patches of 207-2413
#   c2 3.30M points, ~2 500 patches of 52-18 332
#   c4 5.49M points, ~3 400 patches of 22-46 760
SYNTH_PRESETS = {
    # ~120k voxels/frame, 9-bit, ~1000-point patches (~120 per frame), GoP 8, qp 8
    "c0": dict(grid_bits=9, gop=8, qp_i=8, qp_p=8, n_humanoids=1, jitter=0, patch_min=1100, patch_max=2000,
               scale=0.85, shell=0.9, patch_mu=float(np.log(1400)), patch_sigma=0.2),
    # ~800k points/frame, 10-bit, ~800 patches of lognormal 200-3000 points, GoP 8, qp 4, +-1 voxel jitter
    "c1": dict(grid_bits=10, gop=8, qp_i=4, qp_p=4, n_humanoids=1, jitter=1, patch_min=400, patch_max=3000,
               scale=0.85, shell=1.5, patch_mu=float(np.log(1500)), patch_sigma=0.7),
    # ~3M points/frame, 11-bit, heavy-tailed lognormal 50-20000, GoP 16
    "c2": dict(grid_bits=11, gop=16, qp_i=4, qp_p=4, n_humanoids=1, jitter=1, patch_min=80, patch_max=20000,
               scale=0.85, shell=1.5, patch_mu=float(np.log(1500)), patch_sigma=1.6),
    # serving densities: 100k / 9-bit, 400k / 10-bit, 1M / 10-bit
    "c3_100k": dict(grid_bits=9, gop=8, qp_i=4, qp_p=4, n_humanoids=1, jitter=0, patch_min=400, patch_max=3000,
                    scale=0.85, shell=0.75, patch_mu=float(np.log(1500)), patch_sigma=0.7),
    "c3_400k": dict(grid_bits=10, gop=8, qp_i=4, qp_p=4, n_humanoids=1, jitter=0, patch_min=400, patch_max=3000,
                    scale=0.85, shell=0.75, patch_mu=float(np.log(1500)), patch_sigma=0.7),
    "c3_1m": dict(grid_bits=10, gop=8, qp_i=4, qp_p=4, n_humanoids=1, jitter=1, patch_min=400, patch_max=3000,
                  scale=0.85, shell=1.85, patch_mu=float(np.log(1500)), patch_sigma=0.7),
    # 8 humanoids, 12-bit, ~5M points/frame, GoP 32, qp 1, patches 20-50000
    "c4": dict(grid_bits=12, gop=32, qp_i=1, qp_p=1, n_humanoids=8, jitter=1, patch_min=20, patch_max=50000,
               scale=0.85, shell=0.7, patch_mu=float(np.log(1500)), patch_sigma=2.0),
}


def synth_config(preset: str = "c0", n_frames: int | None = None, seed: int = 1, threads: int = 8,
                 zstd_level: int = 3, **over) -> SynthConfig:
    p = dict(SYNTH_PRESETS[preset])
    p.update(over)
    cfg = SynthConfig()
    cfg.grid_bits = p["grid_bits"]
    cfg.n_frames = n_frames if n_frames is not None else p["gop"]
    cfg.gop = p["gop"]
    cfg.qp_i, cfg.qp_p = p["qp_i"], p["qp_p"]
    cfg.n_humanoids = p["n_humanoids"]
    cfg.jitter = p["jitter"]
    cfg.patch_min, cfg.patch_max = p["patch_min"], p["patch_max"]
    cfg.zstd_level = zstd_level
    cfg.threads = threads
    cfg.scale, cfg.shell = p["scale"], p["shell"]
    cfg.patch_mu, cfg.patch_sigma = p["patch_mu"], p["patch_sigma"]
    cfg.noise_sigma = p.get("noise_sigma", 4.0)
    cfg.max_rot_deg = p.get("max_rot_deg", 2.0)
    cfg.max_trans = p.get("max_trans", 2.0)
    cfg.seed = seed
    return cfg


def synth_stream(preset: str = "c0", n_frames: int | None = None, seed: int = 1, threads: int = 8,
                 zstd_level: int = 3, keep_source: bool = False, **over):
    """Generates a seeded synthetic stream; returns the .frame images, or,
    with keep_source, (images, sources) where sources[i] = (xyz f32 [n,3],
    rgb u8 [n,3]) is the cloud frame i was encoded from."""
    L = enc_lib()
    cfg = synth_config(preset, n_frames, seed, threads, zstd_level, **over)
    cfg.keep_source = int(bool(keep_source))
    h = C.c_void_p()
    _check(L.pvvc_synth_generate(C.byref(cfg), C.byref(h)), "pvvc_synth_generate")
    try:
        out, src = [], []
        for i in range(L.pvvc_stream_frames(h)):
            n = C.c_size_t()
            p = L.pvvc_stream_frame_bytes(h, i, C.byref(n))
            out.append(C.string_at(p, n.value))
            if keep_source:
                xp, rp = C.c_void_p(), C.c_void_p()
                m = L.pvvc_stream_frame_source(h, i, C.byref(xp), C.byref(rp))
                xyz = np.ctypeslib.as_array(C.cast(xp, C.POINTER(C.c_float)), (m, 3)).copy() if m else \
                    np.zeros((0, 3), np.float32)
                rgb = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_uint8)), (m, 3)).copy() if m else \
                    np.zeros((0, 3), np.uint8)
                src.append((xyz, rgb))
        return (out, src) if keep_source else out
    finally:
        L.pvvc_stream_free(h)


from . import quality  # noqa: E402,F401  (MSE / PSNR; imports DeviceBuffer lazily)
