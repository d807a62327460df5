"""ctypes bindings of the package's native libraries (include/pvvc_hip.h).

libpvvc_hip.so holds the HIP decoder; libpvvc_enc.so the stream writer.
Loading fails loudly when a library is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

_LIB_DIR = Path(__file__).resolve().parent / "_lib"

c_u8p = C.POINTER(C.c_uint8)
c_i32p = C.POINTER(C.c_int)
c_u32p = C.POINTER(C.c_uint32)
c_f32p = C.POINTER(C.c_float)
c_i64p = C.POINTER(C.c_int64)
c_vp = C.c_void_p


class Points(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("x", "y", "z", "r", "g", "b")]


class DropinFrame(C.Structure):
    """pvvc_dropin_frame: device arrays of one launch_cudaProcess call."""
    _fields_ = [
        ("inner_offset", c_vp), ("index", c_vp), ("type", c_vp), ("mv", c_vp), ("size", c_vp),
        ("qp", c_vp), ("geometry", c_vp), ("geometry_size", c_vp), ("color", c_vp), ("color_size", c_vp),
        ("slice_cnt", C.c_int), ("point_cnt", C.c_int64), ("blob", c_vp), ("blob_bytes", C.c_size_t),
    ]


class SliceDesc(C.Structure):
    """pvvc_slice_desc: one .slice file's fields (mv row-major)."""
    _fields_ = [("type", C.c_uint8), ("timestamp", C.c_int32), ("index", C.c_int32), ("size", C.c_uint64),
                ("qp", C.c_uint8), ("mv", C.c_float * 16), ("geometry", c_vp), ("geometry_size", C.c_uint64),
                ("color", c_vp), ("color_size", C.c_uint64)]


class Tuning(C.Structure):
    _fields_ = [("waves_per_patch", C.c_int), ("lds_bytes", C.c_int), ("wgs_per_cu", C.c_int),
                ("queue_chunk", C.c_int), ("priority_items", C.c_int), ("lane_max_points", C.c_int)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("lanes", C.c_int), ("max_slots", C.c_int),
                ("max_points_per_patch", C.c_int), ("out_pool_points", C.c_int64), ("max_batch_points", C.c_int64),
                ("tuning", Tuning), ("lanes_spin", C.c_int)]


class Completion(C.Structure):
    _fields_ = [("stream_id", C.c_int), ("status", C.c_int), ("seq", C.c_int64), ("tag", C.c_uint64),
                ("points", C.c_int64), ("xyz", C.c_void_p), ("rgb", C.c_void_p), ("out_slot", C.c_int),
                ("arrival_s", C.c_double), ("launch_s", C.c_double), ("done_s", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("frames_done", C.c_int64), ("points_done", C.c_int64), ("batches", C.c_int64),
                ("hbm_budget", C.c_uint64), ("hbm_used", C.c_uint64), ("streams_active", C.c_int),
                ("streams_waiting", C.c_int), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("max_ms", C.c_double), ("mpts_per_s", C.c_double), ("host_ms_per_batch", C.c_double),
                ("backlog_frames", C.c_int64)]


class ColorJob(C.Structure):
    _fields_ = [("geometry", c_vp), ("geometry_size", C.c_uint32), ("yuv", c_vp), ("n", C.c_uint32),
                ("qp", C.c_uint8), ("out", c_vp), ("out_cap", C.c_uint32), ("out_size", C.c_uint32),
                ("status", C.c_int)]


class Mse(C.Structure):
    _fields_ = [("p_to_q", C.c_double * 4), ("q_to_p", C.c_double * 4)]


class SynthConfig(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("grid_bits", "n_frames", "gop", "qp_i", "qp_p", "n_humanoids", "jitter",
                                       "patch_min", "patch_max", "zstd_level", "threads")] + \
               [(n, C.c_double) for n in ("scale", "shell", "patch_mu", "patch_sigma", "noise_sigma", "max_rot_deg",
                                          "max_trans")] + [("seed", C.c_uint64), ("keep_source", C.c_int)]


def _sig(lib, name, restype, *argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)


_hip = None
_enc = None


def hip_lib() -> C.CDLL:
    """The HIP decoder library; raises if it has not been built."""
    global _hip
    if _hip is not None:
        return _hip
    path = _LIB_DIR / "libpvvc_hip.so"
    if not path.exists():
        raise RuntimeError(f"HIP decoder library missing: {path} (run patchvvc-plus_amd/build.py)")
    L = C.CDLL(str(path))
    _sig(L, "pvvc_load_frame", C.c_int, C.c_char_p, C.POINTER(c_vp))
    _sig(L, "pvvc_parse_frame", C.c_int, c_vp, C.c_size_t, C.POINTER(c_vp))
    _sig(L, "pvvc_save_slice", C.c_int64, C.c_char_p, C.POINTER(SliceDesc))
    _sig(L, "pvvc_load_slices", C.c_int, C.POINTER(C.c_char_p), C.c_int, C.POINTER(c_vp))
    _sig(L, "pvvc_slices_to_frame_file", C.c_int, C.POINTER(C.c_char_p), C.c_int, C.c_char_p)
    _sig(L, "pvvc_change_slice_to_frame", C.c_int, C.c_char_p, C.c_char_p)
    _sig(L, "pvvc_frame_free", None, c_vp)
    _sig(L, "pvvc_set_host_threads", C.c_int, C.c_int)
    _sig(L, "pvvc_frame_timestamp", C.c_int, c_vp)
    _sig(L, "pvvc_frame_slice_cnt", C.c_uint32, c_vp)
    _sig(L, "pvvc_frame_points", C.c_uint64, c_vp)
    _sig(L, "pvvc_frame_meta", None, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp)
    _sig(L, "pvvc_frame_geometry", c_vp, c_vp, C.c_uint32)
    _sig(L, "pvvc_frame_color", c_vp, c_vp, C.c_uint32)
    _sig(L, "pvvc_state_alloc", c_vp, C.c_int, C.c_int)
    _sig(L, "pvvc_state_free", None, c_vp)
    _sig(L, "pvvc_state_patch_status", C.c_int, c_vp, C.c_int)
    _sig(L, "pvvc_state_set_stream", C.c_int, c_vp, c_vp)
    _sig(L, "pvvc_state_set_dropin_tuning", C.c_int, c_vp, C.c_int, C.c_int)
    _sig(L, "pvvc_state_set_white_small_patches", C.c_int, c_vp, C.c_int)
    _sig(L, "pvvc_state_synchronize", C.c_int, c_vp)
    _sig(L, "pvvc_decoder_set_stream", C.c_int, c_vp, c_vp)
    _sig(L, "pvvc_decoder_set_white_small_patches", C.c_int, c_vp, C.c_int)
    _sig(L, "pvvc_stream_create", C.c_int, C.POINTER(c_vp))
    _sig(L, "pvvc_stream_destroy", C.c_int, c_vp)
    _sig(L, "pvvc_stream_synchronize", C.c_int, c_vp)
    _sig(L, "launch_cudaProcess", None, C.c_int, C.c_int, c_vp, C.c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
         c_vp, c_vp, c_vp, c_vp, C.c_int)
    _sig(L, "pvvc_set_stream", C.c_int, c_vp)
    _sig(L, "pvvc_last_error", C.c_int)
    _sig(L, "pvvc_set_white_small_patches", C.c_int, C.c_int)
    _sig(L, "pvvc_set_dropin_tuning", C.c_int, C.c_int, C.c_int)
    _sig(L, "pvvc_dropin_stage", C.c_int, c_vp, C.POINTER(DropinFrame))
    _sig(L, "pvvc_dropin_release", None, C.POINTER(DropinFrame))
    _sig(L, "pvvc_dropin_stage_bytes", C.c_size_t, c_vp)
    _sig(L, "pvvc_dropin_stage_to", C.c_int, c_vp, c_vp, C.c_size_t, c_vp, C.c_size_t, C.POINTER(DropinFrame))
    _sig(L, "pvvc_host_alloc", c_vp, C.c_size_t)
    _sig(L, "pvvc_host_free", None, c_vp)
    _sig(L, "pvvc_decoder_create", C.c_int, C.POINTER(c_vp), C.c_int, C.c_int)
    _sig(L, "pvvc_decoder_destroy", C.c_int, c_vp)
    _sig(L, "pvvc_decoder_set_tuning", C.c_int, c_vp, C.POINTER(Tuning))
    _sig(L, "pvvc_decoder_get_tuning", C.c_int, c_vp, C.POINTER(Tuning))
    _sig(L, "pvvc_decoder_reset", C.c_int, c_vp)
    _sig(L, "pvvc_decoder_autotune", C.c_int, c_vp, c_vp, C.c_int, C.c_int, c_vp, c_vp, c_vp, C.c_int,
         C.POINTER(Tuning), C.POINTER(C.c_double))
    _sig(L, "pvvc_batch_create", C.c_int, c_vp, C.POINTER(c_vp), C.c_int, C.POINTER(c_vp))
    _sig(L, "pvvc_batch_free", None, c_vp)
    _sig(L, "pvvc_batch_frames", C.c_int, c_vp)
    _sig(L, "pvvc_batch_items", C.c_int, c_vp)
    _sig(L, "pvvc_batch_frame_points", C.c_int64, c_vp, C.c_int)
    _sig(L, "pvvc_batch_frame_offset", C.c_int64, c_vp, C.c_int)
    _sig(L, "pvvc_batch_input_bytes", C.c_int64, c_vp)
    _sig(L, "pvvc_batch_state_bytes", C.c_int64, c_vp)
    _sig(L, "pvvc_batch_status", C.c_int, c_vp, c_vp, C.c_int)
    _sig(L, "pvvc_decode_batch", C.c_int, c_vp, c_vp, C.c_int, C.c_int, c_vp, c_vp, c_vp)
    _sig(L, "pvvc_decode_batch_noalloc", C.c_int, c_vp, c_vp, C.c_int, c_vp, c_vp, c_vp)
    _sig(L, "pvvc_decode_frame", C.c_int, c_vp, c_vp, C.c_int, c_vp, c_vp, c_vp, c_vp, C.c_int)
    _sig(L, "pvvc_decode_gop", C.c_int, c_vp, c_vp, C.c_int, C.c_int, c_vp, c_vp, c_vp, c_vp, C.c_int)
    _sig(L, "pvvc_encode_colors", C.c_int, c_vp, C.c_int, C.POINTER(C.c_double))
    _sig(L, "pvvc_encoder_release", None)
    _sig(L, "pvvc_cloud_mse", C.c_int, c_vp, c_vp, C.c_uint64, c_vp, c_vp, C.c_uint64, c_vp, c_vp, C.POINTER(Mse),
         C.POINTER(C.c_double))
    _sig(L, "pvvc_psnr", C.c_float, C.c_double, C.c_double, C.c_float)
    _sig(L, "pvvc_server_clock", C.c_double)
    _sig(L, "pvvc_server_create", C.c_int, C.POINTER(ServerConfig), C.POINTER(c_vp))
    _sig(L, "pvvc_server_destroy", C.c_int, c_vp)
    _sig(L, "pvvc_server_open_stream", C.c_int, c_vp, C.c_int, C.c_int, C.c_int64, C.c_int, c_i32p)
    _sig(L, "pvvc_server_stream_state", C.c_int, c_vp, C.c_int)
    _sig(L, "pvvc_server_submit", C.c_int, c_vp, C.c_int, c_vp, C.c_double, C.c_uint64)
    _sig(L, "pvvc_server_close_stream", C.c_int, c_vp, C.c_int)
    _sig(L, "pvvc_server_poll", C.c_int, c_vp, C.POINTER(Completion), C.c_int, C.c_int)
    _sig(L, "pvvc_server_release", C.c_int, c_vp, C.c_int, C.c_int)
    _sig(L, "pvvc_server_get_stats", C.c_int, c_vp, C.POINTER(ServerStats))
    _sig(L, "pvvc_server_reset_stats", C.c_int, c_vp)
    _sig(L, "pvvc_dev_alloc", c_vp, C.c_size_t)
    _sig(L, "pvvc_dev_free", None, c_vp)
    _sig(L, "pvvc_memcpy_h2d", C.c_int, c_vp, c_vp, C.c_size_t)
    _sig(L, "pvvc_memcpy_d2h", C.c_int, c_vp, c_vp, C.c_size_t)
    _sig(L, "pvvc_memset", C.c_int, c_vp, C.c_int, C.c_size_t)
    _sig(L, "pvvc_memcpy_d2d", C.c_int, c_vp, c_vp, C.c_size_t)
    _sig(L, "pvvc_synchronize", C.c_int)
    _sig(L, "pvvc_set_timing", C.c_int, C.c_int)
    _sig(L, "pvvc_last_kernel_ms", C.c_double)
    _sig(L, "pvvc_last_launch_times", C.c_int, c_vp, C.c_int)
    _sig(L, "pvvc_last_decode_kernel_times", C.c_int, c_vp, C.c_int)
    _sig(L, "pvvc_device_info", C.c_int, c_i32p, c_i32p, c_i32p, C.c_char_p, C.c_int)
    _hip = L
    return L


def enc_lib() -> C.CDLL:
    """The stream-writer library (host only)."""
    global _enc
    if _enc is not None:
        return _enc
    path = _LIB_DIR / "libpvvc_enc.so"
    if not path.exists():
        raise RuntimeError(f"stream writer library missing: {path} (run patchvvc-plus_amd/build.py)")
    L = C.CDLL(str(path))
    _sig(L, "pvvc_synth_generate", C.c_int, C.POINTER(SynthConfig), C.POINTER(c_vp))
    _sig(L, "pvvc_stream_frames", C.c_int, c_vp)
    _sig(L, "pvvc_stream_frame_bytes", c_vp, c_vp, C.c_int, C.POINTER(C.c_size_t))
    _sig(L, "pvvc_stream_frame_points", C.c_uint64, c_vp, C.c_int)
    _sig(L, "pvvc_stream_frame_source", C.c_uint64, c_vp, C.c_int, C.POINTER(c_vp), C.POINTER(c_vp))
    _sig(L, "pvvc_stream_write_frame", C.c_int, c_vp, C.c_int, C.c_char_p)
    _sig(L, "pvvc_stream_free", None, c_vp)
    _sig(L, "pvvc_enc_octree", C.c_int64, c_vp, C.c_int, c_vp, C.c_int64, c_i32p)
    _sig(L, "pvvc_enc_raht", C.c_int64, c_vp, C.c_int, c_vp, c_vp)
    _sig(L, "pvvc_enc_rlgr", C.c_int64, c_vp, C.c_int64, c_vp, C.c_int64)
    _sig(L, "pvvc_enc_colors_on_tree", C.c_int64, c_vp, C.c_int64, c_vp, C.c_int, C.c_int, c_vp, C.c_int64)
    _sig(L, "pvvc_enc_patch_colors", C.c_int64, c_vp, C.c_int, c_vp, C.c_int, c_vp, C.c_int64, c_i64p, c_vp, c_i32p,
         c_vp, C.c_int64)
    _sig(L, "pvvc_enc_gop_patch", C.c_int64, c_vp, C.c_int, c_vp, C.c_int, C.c_int, C.c_int, C.c_int, c_vp,
         C.c_int64)
    _enc = L
    return L
