"""Host-side checks that need no GPU: the .frame container, the C ABI
surface of libpvvc_hip.so, and the stream writer."""
import re
import struct
from pathlib import Path

import numpy as np
import pytest
from _frames import pack_frame as _pack_frame

REPO = Path(__file__).resolve().parent.parent


def _zstd(data: bytes) -> bytes:
    import ctypes as C
    lib = C.CDLL("libzstd.so.1")
    lib.ZSTD_compressBound.restype = C.c_size_t
    lib.ZSTD_compress.restype = C.c_size_t
    lib.ZSTD_compress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int]
    cap = lib.ZSTD_compressBound(C.c_size_t(len(data)))
    buf = C.create_string_buffer(cap)
    n = lib.ZSTD_compress(buf, cap, data, len(data), 3)
    return buf.raw[:n]


@pytest.fixture
def fixture_frame():
    mv0 = [1, 0, 0, 5, 0, 1, 0, -2, 0, 0, 1, 0.5, 0, 0, 0, 1]
    geo_a = bytes(range(40))
    col_a = bytes([7] * 33)
    col_b = bytes(range(200, 250))
    slices = [
        dict(index=3, type=0x01, size=11, qp=4, mv=mv0, geo=geo_a, col=col_a),
        dict(index=9, type=0x03, size=22, qp=8, mv=list(range(16)), geo=b"", col=col_b),
        # zstd-wrapped geometry and colour (type bits 3 and 4)
        dict(index=12, type=0x01 | 0x08 | 0x10, size=5, qp=1, mv=mv0, geo=_zstd(geo_a * 3), col=_zstd(col_b * 2)),
    ]
    return _pack_frame(42, slices), slices, geo_a, col_b


def test_container_parse_matches_fields(pvvc, oracle, fixture_frame, tmp_path):
    img, slices, geo_a, col_b = fixture_frame
    p = tmp_path / "fix_0.frame"
    p.write_bytes(img)
    f = pvvc.Frame.load(str(p))
    of = oracle.OracleFrame(str(p))
    assert of.rc == 0
    m, om = f.meta(), of.meta()
    assert f.timestamp == 42 == om["timestamp"]
    for k in ("index", "type", "size", "qp", "mv", "geometry_size", "color_size"):
        np.testing.assert_array_equal(m[k], om[k], err_msg=k)
    assert m["index"].tolist() == [3, 9, 12]
    assert m["geometry_size"].tolist() == [40, 0, 120]     # unwrapped sizes
    assert m["color_size"].tolist() == [33, 50, 100]
    np.testing.assert_array_equal(m["mv"][0], np.array(slices[0]["mv"], np.float32))
    assert f.geometry(2) == geo_a * 3 == of.geometry(2)
    assert f.color(2) == col_b * 2 == of.color(2)
    assert f.color(1) == col_b == of.color(1)
    assert f.points == 38


def test_container_error_codes(pvvc, fixture_frame, tmp_path):
    img, *_ = fixture_frame
    L = pvvc.hip_lib()
    import ctypes as C
    h = C.c_void_p()
    assert L.pvvc_load_frame(str(tmp_path / "x.slice").encode(), C.byref(h)) == 1    # wrong name
    assert L.pvvc_load_frame(str(tmp_path / "missing.frame").encode(), C.byref(h)) == 2
    for cut in (3, 20, len(img) - 1):
        q = tmp_path / f"cut{cut}.frame"
        q.write_bytes(img[:cut])
        assert L.pvvc_load_frame(str(q).encode(), C.byref(h)) == 5
    bad = bytearray(img)
    # corrupt the zstd magic of slice 2's geometry blob -> code 6
    pos = img.index(_zstd(bytes(range(40)) * 3)[:4])
    bad[pos] ^= 0xFF
    q = tmp_path / "badz.frame"
    q.write_bytes(bytes(bad))
    assert L.pvvc_load_frame(str(q).encode(), C.byref(h)) == 6


def test_synthetic_stream_parses_identically(pvvc, oracle, c0_files):
    for p in c0_files[:3]:
        f = pvvc.Frame.load(str(p))
        of = oracle.OracleFrame(str(p))
        m, om = f.meta(), of.meta()
        for k in ("index", "type", "size", "qp", "mv", "geometry_size", "color_size"):
            np.testing.assert_array_equal(m[k], om[k])
        for s in range(0, f.slice_cnt, 17):
            assert f.geometry(s) == of.geometry(s)
            assert f.color(s) == of.color(s)


def test_synthetic_stream_is_deterministic(pvvc):
    a = pvvc.synth_stream("c0", n_frames=2, seed=5, threads=4)
    b = pvvc.synth_stream("c0", n_frames=2, seed=5, threads=2)
    assert a == b
    c = pvvc.synth_stream("c0", n_frames=2, seed=6, threads=4)
    assert a != c


def test_synthetic_stream_decodes_to_the_humanoid(oracle, c0_files):
    """Oracle decode of the generated GoP: every slice OK, the intra frame's
    colours close to the (noisy) field, P frames close to I frames moved."""
    dec = oracle.StreamDecoder(fp64=True)
    outs = [dec.decode(oracle.OracleFrame(str(p)), 8) for p in c0_files[:2]]
    for o in outs:
        assert (o["status"] == 0).all()
        assert o["saturated"] < 0.01 * o["rgb"].size
    # lossy codec at qp 8: FP64 and FP32 colours agree to float precision
    d = np.abs(outs[1]["rgbf"].astype(np.float64) - outs[1]["rgb64"])
    assert d.max() < 1e-2


def test_abi_exports_every_declared_symbol(pvvc):
    """libpvvc_hip.so exports each function include/pvvc_hip.h declares."""
    hdr = (REPO / "include" / "pvvc_hip.h").read_text()
    body = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = set(re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(pvvc_\w+|launch_cudaProcess)\s*\(", body, flags=re.M))
    assert "launch_cudaProcess" in names and "pvvc_decode_batch" in names
    L = pvvc.hip_lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_library_fails_loudly_without_device(pvvc):
    """Without a GPU the decoder reports an error instead of falling back."""
    import ctypes as C
    try:
        info = pvvc.device_info()
    except pvvc.PvvcError:
        info = None
    if info is not None:
        pytest.skip("a GPU is present")
    with pytest.raises(pvvc.PvvcError):
        pvvc.Decoder(4, 100)


def test_poisson_schedule_properties():
    """The serving driver's arrival schedule: sorted, per-stream frame numbers
    consecutive from 0, mean gap close to 1/fps."""
    import importlib
    serve = importlib.import_module("pvvc.serve")
    ev = serve.poisson_schedule(16, 30.0, 20.0, seed=5)
    ts = [e[0] for e in ev]
    assert ts == sorted(ts) and ts[0] >= 0 and ts[-1] < 20.0
    per = {}
    for t, s, k in ev:
        assert k == per.get(s, -1) + 1
        per[s] = k
    rate = len(ev) / (16 * 20.0)
    assert 27.0 < rate < 33.0


def test_last_error_is_per_thread(pvvc):
    """A failure on one host thread is that thread's pvvc_last_error alone:
    launch_cudaProcess with an unknown state records INVALID_ARG (before any
    device call) on the calling thread, and the main thread reads 0."""
    import threading
    L = pvvc.hip_lib()
    L.pvvc_last_error()
    seen = {}

    def worker():
        L.launch_cudaProcess(1, 1, None, 0, None, None, None, None, None, None, None, None, None, None,
                             C_void(0xdead0), 3)
        seen["worker"] = L.pvvc_last_error()
        seen["worker_again"] = L.pvvc_last_error()

    t = threading.Thread(target=worker)
    t.start()
    t.join()
    assert seen == {"worker": 8, "worker_again": 0}
    assert L.pvvc_last_error() == 0


def C_void(v):
    import ctypes as C
    return C.c_void_p(v)
