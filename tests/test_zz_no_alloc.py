"""pvvc_decode_batch_noalloc through the C ABI against pvvc_decode_batch's
persistent mode: identical bytes and statuses on two calls in a row (a c1
GoP, whose workgroups the scratch cuts, then a smaller c0 batch with
Points output), and a clear CAPACITY error for a batch beyond the scratch."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CAPACITY = 7


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dims(frames):
    metas = [f.meta() for f in frames]
    return max(int(m["index"].max()) for m in metas) + 1, max(int(m["size"].max()) for m in metas)


def _p(buf):
    return C.c_void_p(buf.ptr) if buf is not None else None


def _both(pvvc, dec, b, output):
    L = pvvc.hip_lib()
    ref = dec.outputs(b, output)
    dec.decode(b, pvvc.MODE_PERSISTENT, ref)
    pvvc.synchronize()
    st_ref = b.status().copy()
    got = dec.outputs(b, output)
    rc = L.pvvc_decode_batch_noalloc(dec.handle, b.handle, output, _p(got.xyz), _p(got.rgb), _p(got.pts))
    assert rc == 0
    pvvc.synchronize()
    np.testing.assert_array_equal(b.status(), st_ref)
    assert (st_ref == 0).all()
    n = b.total_points
    if output == pvvc.OUT_NATIVE:
        for x, y, dt, k in ((ref.xyz, got.xyz, np.float32, 3), (ref.rgb, got.rgb, np.uint8, 3)):
            np.testing.assert_array_equal(y.read(dt, k * n), x.read(dt, k * n))
    else:
        np.testing.assert_array_equal(got.pts.read(pvvc.POINTS_DTYPE, n), ref.pts.read(pvvc.POINTS_DTYPE, n))


def test_noalloc_equals_persistent_decode(pvvc, c0_stream):
    big = [pvvc.Frame.parse(b) for b in pvvc.synth_stream("c1", n_frames=8, seed=4, threads=16)]
    small = [pvvc.Frame.parse(b) for b in c0_stream[:2]]
    np_ = max(_dims(big)[0], _dims(small)[0])
    mp = max(_dims(big)[1], _dims(small)[1])
    dec = pvvc.Decoder(np_, mp)
    b1 = dec.batch(big)
    _say("call 1: c1 GoP,", b1.total_points, "points,", b1.items, "slices")
    _both(pvvc, dec, b1, pvvc.OUT_NATIVE)
    b2 = dec.batch(small)
    _say("call 2: c0 frames 0-1,", b2.total_points, "points")
    _both(pvvc, dec, b2, pvvc.OUT_POINTS)
    _both(pvvc, dec, b2, pvvc.OUT_NATIVE)
    # beyond the scratch: 30 GoPs of c0 in one batch need ~350 MB of stage
    c0 = [pvvc.Frame.parse(b) for b in c0_stream]
    b3 = dec.batch(c0 * 30)
    out = dec.outputs(b3, pvvc.OUT_NATIVE)
    L = pvvc.hip_lib()
    L.pvvc_last_error()
    rc = L.pvvc_decode_batch_noalloc(dec.handle, b3.handle, pvvc.OUT_NATIVE, _p(out.xyz), _p(out.rgb), None)
    assert rc == CAPACITY
    assert L.pvvc_last_error() == CAPACITY
    # still usable afterwards
    _both(pvvc, dec, b2, pvvc.OUT_NATIVE)
