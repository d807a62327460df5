"""Quality metrics: the symmetric nearest-neighbour MSE of vvc::common::MSE
(include/registration/registration.h:48-100,
src/registration/registration_base.cc:159-238) computed on the GPU by
pvvc_cloud_mse, and the PSNR that PVVCDecompression prints from it
(src/codec/pvvc_decompression.cc:150-164).

    m = MSE()
    m.set_clouds((src_xyz, src_rgb), (dec_xyz, dec_rgb))
    m.compute()
    geo = max(*m.get_geo_mses()); y = max(*m.get_y_mses())

A cloud is (xyz float32 [n, 3], rgb uint8 [n, 3]) as numpy arrays (copied
to the device) or as DeviceBuffers already holding that layout, given with
their point count: (xyz_buf, rgb_buf, n).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._native import Mse, hip_lib

GEO_PEAK = 1024.0  # pvvc_decompression.cc:158
Y_PEAK = 256.0  # pvvc_decompression.cc:159


def _device_cloud(cloud, keep: list):
    from . import DeviceBuffer
    if len(cloud) == 3:  # (xyz DeviceBuffer, rgb DeviceBuffer, n)
        xb, cb, n = cloud
        return xb.ptr, cb.ptr, int(n)
    xyz = np.ascontiguousarray(cloud[0], np.float32).reshape(-1, 3)
    rgb = np.ascontiguousarray(cloud[1], np.uint8).reshape(-1, 3)
    if len(xyz) != len(rgb):
        raise ValueError("xyz and rgb differ in point count")
    n = len(xyz)
    xb, cb = DeviceBuffer(max(1, xyz.nbytes)), DeviceBuffer(max(1, rgb.nbytes))
    if n:
        xb.write(xyz)
        cb.write(rgb)
    keep += [xb, cb]
    return xb.ptr, cb.ptr, n


def cloud_mse(p, q, want_nn: bool = False) -> dict:
    """MSE::Compute on the GPU.  Returns geo/y/u/v -> (p->q, q->p) as FP64
    means, 'kernel_ms', and with want_nn 'nn_p' / 'nn_q' (int64 neighbour
    index of every point).  Ties go to the lowest index."""
    from . import DeviceBuffer, PvvcError
    keep: list = []
    pxp, pcp, npnt = _device_cloud(p, keep)
    qxp, qcp, nqnt = _device_cloud(q, keep)
    nnp = DeviceBuffer(8 * max(1, npnt)) if want_nn else None
    nnq = DeviceBuffer(8 * max(1, nqnt)) if want_nn else None
    out = Mse()
    ms = C.c_double()
    rc = hip_lib().pvvc_cloud_mse(C.c_void_p(pxp), C.c_void_p(pcp), npnt, C.c_void_p(qxp), C.c_void_p(qcp), nqnt,
                                  C.c_void_p(nnp.ptr if nnp else None), C.c_void_p(nnq.ptr if nnq else None),
                                  C.byref(out), C.byref(ms))
    if rc != 0:
        raise PvvcError(rc, "pvvc_cloud_mse")
    res = {k: (float(out.p_to_q[i]), float(out.q_to_p[i])) for i, k in enumerate(("geo", "y", "u", "v"))}
    res["kernel_ms"] = float(ms.value)
    if want_nn:
        res["nn_p"] = nnp.read(np.int64, npnt)
        res["nn_q"] = nnq.read(np.int64, nqnt)
    return res


def psnr(mse_pair, peak: float) -> float:
    """10 log10(peak^2 / max(pair)) in FP32, as pvvc_decompression.cc:156-159."""
    return float(hip_lib().pvvc_psnr(float(mse_pair[0]), float(mse_pair[1]), float(peak)))


class MSE:
    """vvc::common::MSE: set_clouds(p, q), compute(), get_*_mses() -> (p->q, q->p)
    as float32 values like the reference's std::pair<float, float>."""

    def __init__(self):
        self._p = self._q = None
        self._r = {k: (0.0, 0.0) for k in ("geo", "y", "u", "v")}
        self.kernel_ms = 0.0

    def set_clouds(self, p, q) -> None:
        self._p, self._q = p, q

    def compute(self) -> None:
        def empty(c):
            return c is None or (len(c) == 3 and int(c[2]) == 0) or (len(c) == 2 and len(c[0]) == 0)
        if empty(self._p) or empty(self._q):  # registration_base.cc:170-179
            a = -1.0 if empty(self._p) else 0.0
            b = -1.0 if empty(self._q) else 0.0
            self._r = {k: (a, b) for k in self._r}
            return
        r = cloud_mse(self._p, self._q)
        self.kernel_ms = r["kernel_ms"]
        self._r = {k: (float(np.float32(r[k][0])), float(np.float32(r[k][1]))) for k in self._r}

    def get_geo_mses(self) -> tuple[float, float]:
        return self._r["geo"]

    def get_y_mses(self) -> tuple[float, float]:
        return self._r["y"]

    def get_u_mses(self) -> tuple[float, float]:
        return self._r["u"]

    def get_v_mses(self) -> tuple[float, float]:
        return self._r["v"]
