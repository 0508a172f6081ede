"""On-demand drum metrics on the GPU over ``include/millgpu.h`` (crates/mill-core/src/metrics.rs).

``mixing_index(points, dye, radius)`` mirrors ``metrics::mixing_index`` (metrics.rs:506-576) and
returns None where the reference does; ``mixing_index_batch`` evaluates many drums in one launch.
Inputs are copied into device memory of the library's own HIP runtime; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import sim as _sim
from .grid import DeviceArray, _lib as _grid_lib

_ready = set()


def _lib():
    L = _grid_lib()
    if id(L) in _ready:
        return L
    P = C.c_void_p
    L.mg_mixing_index.argtypes = [P, P, C.c_uint32, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_int)]
    L.mg_mixing_index_batch.argtypes = [P, P, P, P, C.c_uint32, P, P, P]
    _ready.add(id(L))
    return L


def mixing_index(points, dye, radius: float):
    """metrics::mixing_index of one drum's fluid: the Lacey index, or None."""
    L = _lib()
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float32).reshape(-1, 2))
    d = np.ascontiguousarray(np.asarray(dye, dtype=np.float32).reshape(-1))
    if d.size != pts.shape[0]:
        raise ValueError("one dye value per particle")
    dx, dd = DeviceArray.from_numpy(pts), DeviceArray.from_numpy(d)
    out, ok = C.c_float(), C.c_int()
    _sim._check(L.mg_mixing_index(dx.ptr, dd.ptr, d.size, float(radius), C.byref(out), C.byref(ok)),
                what="mg_mixing_index")
    return float(out.value) if ok.value else None


def mixing_index_batch(drums):
    """Mixing indices of several drums in one launch: drums = [(points, dye, radius), ...]."""
    L = _lib()
    pts = [np.asarray(p, dtype=np.float32).reshape(-1, 2) for p, _, _ in drums]
    dye = [np.asarray(d, dtype=np.float32).reshape(-1) for _, d, _ in drums]
    off = np.zeros(len(drums) + 1, dtype=np.uint32)
    off[1:] = np.cumsum([len(d) for d in dye])
    rad = np.array([r for _, _, r in drums], dtype=np.float32)
    allp = np.concatenate(pts) if pts else np.zeros((0, 2), np.float32)
    alld = np.concatenate(dye) if dye else np.zeros(0, np.float32)
    dx, dd, doff, drad = (DeviceArray.from_numpy(a) for a in (allp, alld, off, rad))
    dout, dok = DeviceArray(4 * len(drums)), DeviceArray(4 * len(drums))
    _sim._check(L.mg_mixing_index_batch(dx.ptr, dd.ptr, doff.ptr, drad.ptr, len(drums), dout.ptr, dok.ptr, None),
                what="mg_mixing_index_batch")
    out = dout.to_numpy(np.float32, len(drums))
    ok = dok.to_numpy(np.int32, len(drums))
    return [float(v) if k else None for v, k in zip(out, ok)]
