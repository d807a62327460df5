"""The .slice per-patch files (io::SaveSlice / io::LoadSlice,
src/io/slice_io.cc:20-206) and their assembly into a frame
(io::ChangeSliceToFrame, slice_io.cc:208-315; Frame(slices),
include/common/common.h:378-412), product against the oracle's
restatement and against hand-packed bytes.

* A hand-packed .slice (a non-symmetric motion matrix, stored column-major
  as Eigen::Matrix4f is) equals what the product and the oracle write.
* Slices in arbitrary file order become the frame image packed by hand from
  the index-sorted slices: the first sorted slice's timestamp, u32 sizes,
  the motion row-major, a predictive slice without geometry.
* Synthetic c0 frames (an intra frame with zstd-wrapped blobs, a
  predictive frame) split into .slice files come back as the same .frame
  bytes, and pvvc_load_slices parses to the same frame.
* Error codes of both, including a declared blob larger than the file.
"""
import struct

import numpy as np
import pytest

from _frames import pack_frame

MV = np.array([[0.0, -1.0, 0.0, 5.5], [1.0, 0.0, 0.0, -2.0], [0.0, 0.0, 2.0, 0.25], [0.0, 0.0, 0.0, 1.0]],
              np.float32)  # rows


def pack_slice(type_, ts, index, size, qp, mv_rows, geo=None, col=None):
    """The SaveSlice layout, field by field (slice_io.cc:44-116)."""
    out = struct.pack("<BiiQB", type_, ts, index, size, qp)
    out += np.asarray(mv_rows, np.float32).T.astype("<f4").tobytes()  # column-major storage
    if not type_ & 0x02:
        out += struct.pack("<Q", len(geo)) + geo
    if not type_ & 0x04:
        out += struct.pack("<Q", len(col)) + col
    return out


def split_frame(image):
    """The raw fields of a .frame image (blobs as stored, zstd included)."""
    ts, S = struct.unpack_from("<iI", image, 0)
    o = 8
    index = struct.unpack_from(f"<{S}i", image, o); o += 4 * S
    types = image[o:o + S]; o += S
    size = struct.unpack_from(f"<{S}I", image, o); o += 4 * S
    qp = image[o:o + S]; o += S
    mv = np.frombuffer(image, "<f4", 16 * S, o).reshape(S, 4, 4); o += 64 * S
    gs = struct.unpack_from(f"<{S}I", image, o); o += 4 * S
    geo = []
    for g in gs:
        geo.append(image[o:o + g]); o += g
    cs = struct.unpack_from(f"<{S}I", image, o); o += 4 * S
    col = []
    for c in cs:
        col.append(image[o:o + c]); o += c
    assert o == len(image)
    return ts, [dict(index=index[i], type=types[i], size=size[i], qp=qp[i], mv=mv[i], geo=geo[i], col=col[i])
                for i in range(S)]


def test_slice_bytes_hand_product_oracle(pvvc, oracle, tmp_path):
    geo, col = bytes(range(30)), b"\x11\x22\x33"
    want = pack_slice(0x01, 40, 7, 123, 4, MV, geo, col)
    p1, p2 = tmp_path / "a.slice", tmp_path / "b.slice"
    assert pvvc.save_slice(p1, 0x01, 40, 7, 123, 4, MV, geo, col) == len(want)
    assert oracle.save_slice(p2, 0x01, 40, 7, 123, 4, MV, geo, col) == len(want)
    assert p1.read_bytes() == want == p2.read_bytes()
    # predictive: no geometry field at all
    want = pack_slice(0x03, 41, 7, 123, 4, MV, None, col)
    assert pvvc.save_slice(p1, 0x03, 41, 7, 123, 4, MV, None, col) == len(want)
    assert p1.read_bytes() == want


def test_slices_to_frame_hand_packed(pvvc, oracle, tmp_path):
    rng = np.random.default_rng(3)
    recs = [dict(index=5, type=0x01, ts=9, size=300, qp=4, geo=rng.bytes(40), col=rng.bytes(17)),
            dict(index=1, type=0x03, ts=8, size=50, qp=2, geo=None, col=rng.bytes(9)),
            dict(index=3, type=0x01, ts=9, size=70000, qp=8, geo=rng.bytes(5), col=rng.bytes(3))]
    paths = []
    for k, r in enumerate(recs):
        mv = MV + k
        r["mv"] = mv
        p = tmp_path / f"s{k}.slice"
        p.write_bytes(pack_slice(r["type"], r["ts"], r["index"], r["size"], r["qp"], mv, r["geo"], r["col"]))
        paths.append(p)
    srt = sorted(recs, key=lambda r: r["index"])
    want = pack_frame(srt[0]["ts"], [dict(index=r["index"], type=r["type"], size=r["size"], qp=r["qp"],
                                          mv=[float(x) for x in r["mv"].reshape(16)], geo=r["geo"] or b"",
                                          col=r["col"]) for r in srt])
    pf, of = tmp_path / "p.frame", tmp_path / "o.frame"
    pvvc.slices_to_frame_file(paths, pf)
    assert oracle.slices_to_frame(paths, of) == 0
    assert pf.read_bytes() == want == of.read_bytes()
    f = pvvc.Frame.from_slices(paths)
    m = f.meta()
    assert f.timestamp == 8 and list(m["index"]) == [1, 3, 5]
    np.testing.assert_array_equal(m["mv"][0], (MV + 1).reshape(16))  # row-major
    assert f.geometry(0) == b"" and f.color(0) == recs[1]["col"]


@pytest.mark.parametrize("frame", [0, 1], ids=["intra_zstd", "predictive"])
def test_frame_split_into_slices_round_trips(pvvc, oracle, tmp_path, frame):
    image = pvvc.synth_stream("c0", n_frames=2, seed=4, threads=8)[frame]
    ts, slices = split_frame(image)
    if frame == 0:
        assert any(s["type"] & 0x18 for s in slices)  # zstd-wrapped blobs
    else:
        assert all(s["type"] & 0x02 for s in slices)  # predictive: no geometry
    d = tmp_path / "slices"
    d.mkdir()
    paths = []
    for s in reversed(slices):
        p = d / f"seq_time_{ts:04d}_slice_{s['index']:04d}.slice"
        p.write_bytes(pack_slice(s["type"], ts, s["index"], s["size"], s["qp"], s["mv"], s["geo"] or None,
                                 s["col"]))
        paths.append(p)
    (d / "notes_slice.txt").write_bytes(b"not a slice")
    for out, run in ((tmp_path / "p.frame", lambda o: pvvc.slices_to_frame_file(paths, o)),
                     (tmp_path / "d.frame", lambda o: pvvc.change_slice_to_frame(d, o))):
        run(out)
        assert out.read_bytes() == image
    assert oracle.slices_to_frame(paths, tmp_path / "o.frame") == 0
    assert (tmp_path / "o.frame").read_bytes() == image
    a, b = pvvc.Frame.from_slices(paths), pvvc.Frame.parse(image)
    ma, mb = a.meta(), b.meta()
    for k in ma:
        np.testing.assert_array_equal(ma[k], mb[k])
    for i in range(a.slice_cnt):
        assert a.geometry(i) == b.geometry(i) and a.color(i) == b.color(i)


def test_slice_errors(pvvc, oracle, tmp_path):
    geo, col = b"\x01" * 30, b"\x02\x03"
    good = pack_slice(0x01, 1, 2, 3, 4, MV, geo, col)
    cases = {
        "name.slc": (good, 1),
        "missing.slice": (None, 2),
        "novalid.slice": (bytes([0x02]) + good[1:], 7),
        "short.slice": (good[:-1], 5),
        "short_header.slice": (good[:10], 5),
        "zero_geo.slice": (good[:1 + 4 + 4 + 8 + 1 + 64] + struct.pack("<Q", 0), 8),
        "huge_geo.slice": (good[:1 + 4 + 4 + 8 + 1 + 64] + struct.pack("<Q", 1 << 50) + geo, 5),
    }
    for name, (data, code) in cases.items():
        p = tmp_path / name
        if data is not None:
            p.write_bytes(data)
        with pytest.raises(pvvc.PvvcError) as e:
            pvvc.Frame.from_slices([p])
        assert e.value.code == code, name
        assert oracle.slices_to_frame([p], tmp_path / "x.frame") == code, name
    ok = tmp_path / "ok.slice"
    ok.write_bytes(good + b"trailing")  # extra bytes: the reference only warns
    assert pvvc.Frame.from_slices([ok]).slice_cnt == 1
    with pytest.raises(pvvc.PvvcError) as e:
        pvvc.Frame.from_slices([])
    assert e.value.code == 9
    with pytest.raises(pvvc.PvvcError) as e:
        pvvc.slices_to_frame_file([ok], tmp_path / "out.frm")
    assert e.value.code == 1
    assert oracle.slices_to_frame([ok], tmp_path / "out.frm") == 1
    empty = tmp_path / "empty"
    empty.mkdir()
    for d, code in ((empty, 2), (tmp_path / "nodir", 4)):
        with pytest.raises(pvvc.PvvcError) as e:
            pvvc.change_slice_to_frame(d, tmp_path / "x.frame")
        assert e.value.code == code
    with pytest.raises(pvvc.PvvcError) as e:
        pvvc.save_slice(tmp_path / "w.slice", 0x01, 0, 0, 1, 1, MV, None, col)  # intra without geometry
    assert e.value.code == 8
    assert oracle.save_slice(tmp_path / "w.slice", 0x01, 0, 0, 1, 1, MV, None, col) == -8
