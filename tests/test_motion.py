"""Patch motion (pvvc_decompression.cc:141-149; CUDA2OpenGL.cu:44-58):
x' = x*m0 + y*m1 + z*m2 + m3 over rows 0-2 of the row-major 4x4.

Pinned by closed forms: the eight leaves of a full node, (0.5 | 1.5) per
axis, moved by matrices whose entries are small integers or powers of two,
so that every product and sum is exact in FP32 whatever the order.  The
matrices are not symmetric, so a transposed matrix, a wrong row or column
index, a dropped term or a missing translation changes the result."""
import struct

import numpy as np
import pytest

from _frames import pack_frame

MATS = {
    "rot_z_90": [[0, -1, 0, 3], [1, 0, 0, -2], [0, 0, 1, 0.5]],
    "rot_x_scale2": [[2, 0, 0, 1], [0, 0, -2, 3], [0, 2, 0, -5]],
    "perm_mixed": [[0, 0, 0.5, 7], [-1, 0, 0, 0.25], [0, 4, 0, -2]],
    "shear": [[1, 2, 0, 0], [0, 1, -3, 1], [1, 0, 1, 2]],
}
# leaves of header (1,1,1)/(2,2,2)/h=2 with occupancy FF, in decoder order
LEAVES = np.array([(1.5 if not (i & 4) else 0.5, 1.5 if not (i & 2) else 0.5, 1.5 if not (i & 1) else 0.5)
                   for i in range(8)], np.float64)


def _frame(oracle):
    geo = struct.pack(">6f", 1, 1, 1, 2, 2, 2) + bytes([2, 0xFF])
    sym = np.zeros(24, np.int64)
    sym[7], sym[15], sym[23] = 300, 360, 370  # DC of Y, U, V
    col = oracle.rlgr_encode(sym)
    slices, expect = [], []
    for k, (name, m) in enumerate(MATS.items()):
        m = np.array(m, np.float64)
        mv = list(m.reshape(-1)) + [0, 0, 0, 1]
        slices.append(dict(index=k, type=0x01, size=8, qp=1, mv=mv, geo=geo, col=col))
        expect.append(LEAVES @ m[:, :3].T + m[:, 3])
    return pack_frame(0, slices), np.concatenate(expect).astype(np.float32)


def test_motion_closed_forms_oracle(oracle, tmp_path):
    img, want = _frame(oracle)
    p = tmp_path / "motion_0.frame"
    p.write_bytes(img)
    out = oracle.StreamDecoder(fp64=False).decode(oracle.OracleFrame(str(p)), 2)
    assert (out["status"] == 0).all()
    np.testing.assert_array_equal(out["xyz"], want)
    # the closed forms are exact, so they also equal the FP64 products
    assert np.array_equal(want.astype(np.float64), np.concatenate(
        [LEAVES @ np.array(m, np.float64)[:, :3].T + np.array(m, np.float64)[:, 3] for m in MATS.values()]))


def test_motion_transposed_matrix_differs():
    """The closed forms separate a matrix from its transpose (the test would
    catch a row/column mix-up)."""
    for m in MATS.values():
        m = np.array(m, np.float64)
        assert not np.array_equal(LEAVES @ m[:, :3].T, LEAVES @ m[:, :3])


@pytest.mark.gpu
def test_motion_closed_forms_gpu(pvvc, oracle, tmp_path):
    img, want = _frame(oracle)
    f = pvvc.Frame.parse(img)
    for mode in (pvvc.MODE_PER_FRAME, pvvc.MODE_PERSISTENT):
        dec = pvvc.Decoder(len(MATS), 8)
        b = dec.batch([f])
        xyz, _ = dec.decode_host(b, mode, pvvc.OUT_NATIVE)
        assert (b.status() == 0).all()
        np.testing.assert_array_equal(xyz, want)
    st = pvvc.State(len(MATS), 8)
    pts = pvvc.decode_dropin(f, st)
    np.testing.assert_array_equal(np.stack([pts["x"], pts["y"], pts["z"]], 1), want)
