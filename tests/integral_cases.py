"""Inputs for the exact-sum pins of the compensated integrals: fields of more
than 10^6 cells on which the reference's serial sum is off the exactly rounded
sum by many ulps, and the oracle's per-cell addends of each sum (the exports
oracle_integral_terms and oracle_avegradphilocal_terms; the CPU tests tie them
to the oracle's sums, which are bitwise the reference's)."""
import ctypes
import math

import numpy as np

import harness as H

N = (112, 104, 96)  # 1.12e6 interior cells
GB = (-3, N[0] + 2, -3, N[1] + 2, -3, N[2] + 2)
IB = (0, N[0] - 1, 0, N[1] - 1, 0, N[2] - 1)
H_ = 0.1  # dV = 0.1^3 is not a binary fraction: every addend rounds
EPS = 0.15


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _ints(*v):
    cs = [ctypes.c_int(x) for x in v]
    return cs, [ctypes.byref(c) for c in cs]


def volume_case(seed=0):
    """phi: 3/4 of the cells deep inside (each adds dV), the rest spread over
    the smoothing band and outside; gradients for the area."""
    rng = np.random.default_rng(seed)
    shp = H.shape_of(GB)
    phi = -1.0 - rng.random(shp)
    band = rng.random(shp) < 0.25
    phi[band] = EPS * (2.0 * rng.random(int(band.sum())) - 1.0) * 1.2
    grads = [rng.standard_normal(shp) for _ in range(3)]
    return phi, grads


def integral_terms(which, phi, grads=None, gbox=GB, ibox=IB, h=H_, eps=EPS):
    """The oracle's addends of the volume (which 0: phi < 0, 1: phi > 0) or area
    (2) sum over ibox, in loop order (the volumes read no gradients)."""
    lib = ctypes.CDLL(H.ORACLE_SO)
    f = lib.oracle_integral_terms
    f.restype = None
    n = (ibox[1] - ibox[0] + 1) * (ibox[3] - ibox[2] + 1) * (ibox[5] - ibox[4] + 1)
    out = np.empty(n)
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    grads = grads if grads is not None else [phi] * 3
    k1, gb = _ints(*gbox)
    k2, ib = _ints(*ibox)
    d = [ctypes.c_double(v) for v in (h, h, h, eps)]
    f(_ptr(out), ctypes.c_int(which), _ptr(phi), *gb, _ptr(grads[0]), _ptr(grads[1]), _ptr(grads[2]), *gb, *ib,
      *[ctypes.byref(x) for x in d])
    return out


def band_all(phi_shape):
    """An index list over every interior cell (k, j, i order), marks 1."""
    k, j, i = np.meshgrid(np.arange(IB[4], IB[5] + 1), np.arange(IB[2], IB[3] + 1), np.arange(IB[0], IB[1] + 1),
                          indexing="ij")
    nb = np.zeros(phi_shape, np.uint8)
    nb[3:-3, 3:-3, 3:-3] = 1
    return (i.reshape(-1).astype(np.int32), j.reshape(-1).astype(np.int32), k.reshape(-1).astype(np.int32), nb)


def avegrad_case(seed=1):
    rng = np.random.default_rng(seed)
    phi = rng.standard_normal(H.shape_of(GB)) * 0.3
    return phi


def avegrad_terms(phi, ix, iy, iz, nb):
    lib = ctypes.CDLL(H.ORACLE_SO)
    f = lib.oracle_avegradphilocal_terms
    f.restype = None
    out = np.empty(ix.size)
    cnt = ctypes.c_int(0)
    k1, gb = _ints(*GB)
    d = [ctypes.c_double(H_) for _ in range(3)]
    lo, hi = ctypes.c_int(0), ctypes.c_int(ix.size - 1)
    mark = ctypes.c_ubyte(1)
    f(_ptr(out), ctypes.byref(cnt), _ptr(phi), *gb, *[ctypes.byref(x) for x in d], _ptr(ix), _ptr(iy), _ptr(iz),
      ctypes.byref(lo), ctypes.byref(hi), _ptr(nb), *gb, ctypes.byref(mark))
    return out[:cnt.value]


def serial(terms):
    s = 0.0
    for t in terms.tolist():
        s = s + t
    return s


def ulps(a, b):
    return abs(a - b) / math.ulp(abs(b))
