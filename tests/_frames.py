"""Test helpers: .frame images packed field by field in the order of
ChangeSliceToFrame (src/io/slice_io.cc:258-315), and unpacked through the
product parser for mutation."""
import struct


def pack_frame(ts, slices):
    S = len(slices)
    out = struct.pack("<iI", ts, S)
    out += b"".join(struct.pack("<i", s["index"]) for s in slices)
    out += bytes(s["type"] for s in slices)
    out += b"".join(struct.pack("<I", s["size"]) for s in slices)
    out += bytes(s["qp"] for s in slices)
    out += b"".join(struct.pack("<16f", *s["mv"]) for s in slices)
    out += b"".join(struct.pack("<I", len(s["geo"])) for s in slices)
    out += b"".join(s["geo"] for s in slices)
    out += b"".join(struct.pack("<I", len(s["col"])) for s in slices)
    out += b"".join(s["col"] for s in slices)
    return out


def unpack_frame(pvvc, image: bytes):
    """(timestamp, slices) with unwrapped blobs and the zstd type bits cleared."""
    f = pvvc.Frame.parse(image)
    m = f.meta()
    slices = []
    for s in range(f.slice_cnt):
        slices.append(dict(index=int(m["index"][s]), type=int(m["type"][s]) & 0x07, size=int(m["size"][s]),
                           qp=int(m["qp"][s]), mv=[float(x) for x in m["mv"][s]],
                           geo=f.geometry(s), col=f.color(s)))
    return f.timestamp, slices


def bits_to_bytes(bits: str) -> bytes:
    bits = bits + "0" * (-len(bits) % 8)
    return bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8))
