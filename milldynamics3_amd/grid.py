"""The reference's uniform grid (crates/mill-core/src/grid.rs) built on the GPU, over ``include/millgpu.h``.

``UniformGrid.build(points, cell_size)`` mirrors ``UniformGrid::build`` (grid.rs:101-153): same
cell keys, Dense/Sparse layout rule, x-major cell order and ascending point index within a cell.
``candidate_pairs()`` is ``for_each_candidate_pair`` (grid.rs:291-362) as a list in emission order,
``near(queries)`` is ``for_each_near`` (grid.rs:371-399) and ``neighbor_lists(h)`` is
``build_neighbor_lists`` (pbf.rs:262-288). Points are copied into device memory of the library's
own HIP runtime (``DeviceArray``); the compute path is the HIP library and there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import sim as _sim

_ready = set()


def _lib():
    L = _sim.lib()
    if id(L) in _ready:
        return L
    P = C.c_void_p
    u32p = C.POINTER(C.c_uint32)
    u64p = C.POINTER(C.c_uint64)
    L.mg_grid_create.argtypes = [C.c_uint32, C.POINTER(P), C.c_char_p, C.c_size_t]
    L.mg_grid_destroy.argtypes = [P]
    L.mg_grid_build.argtypes = [P, P, C.c_uint32, C.c_float, P]
    L.mg_grid_meta.argtypes = [P, C.POINTER(C.c_int64)]
    L.mg_grid_copy.argtypes = [P, u32p, C.c_size_t, u32p, C.c_size_t, C.POINTER(C.c_int32), C.c_size_t]
    L.mg_grid_candidate_pairs.argtypes = [P, P, C.c_uint64, u64p, P]
    L.mg_grid_neighbor_lists.argtypes = [P, P, C.c_float, P, P, C.c_uint64, u64p, P]
    L.mg_grid_near.argtypes = [P, P, C.c_uint32, P, P, C.c_uint64, u64p, P]
    L.mg_grid_last_error.argtypes = [P]
    L.mg_grid_last_error.restype = C.c_char_p
    L.mg_device_malloc.argtypes = [C.c_size_t, C.POINTER(P)]
    L.mg_device_free.argtypes = [P]
    L.mg_memcpy_htod.argtypes = [P, P, C.c_size_t]
    L.mg_memcpy_dtoh.argtypes = [P, P, C.c_size_t]
    L.mg_device_synchronize.argtypes = []
    _ready.add(id(L))
    return L


def synchronize():
    """Waits for all work of the library's own HIP runtime on its current device (the library's
    streams are not torch's: ``torch.cuda.synchronize()`` does not wait for them)."""
    _sim._check(_lib().mg_device_synchronize(), what="mg_device_synchronize")


class DeviceArray:
    """Device memory of the library's own HIP runtime (numpy in, numpy out)."""

    def __init__(self, nbytes: int):
        self._L = _lib()
        self.ptr = C.c_void_p()
        self.nbytes = int(nbytes)
        _sim._check(self._L.mg_device_malloc(max(self.nbytes, 4), C.byref(self.ptr)), what="mg_device_malloc")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        _sim._check(d._L.mg_memcpy_htod(d.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes), what="mg_memcpy_htod")
        return d

    def to_numpy(self, dtype, count: int):
        out = np.zeros(int(count), dtype=dtype)
        _sim._check(self._L.mg_memcpy_dtoh(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes), what="mg_memcpy_dtoh")
        return out

    def free(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            self._L.mg_device_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        self.free()


def _points(points):
    return np.ascontiguousarray(np.asarray(points, dtype=np.float32).reshape(-1, 2))


class UniformGrid:
    """A grid handle with device workspace for up to ``max_points`` points."""

    def __init__(self, max_points: int):
        self._L = _lib()
        self._h = C.c_void_p()
        err = C.create_string_buffer(256)
        _sim._check(self._L.mg_grid_create(int(max_points), C.byref(self._h), err, 256), err, "mg_grid_create")
        self.points = None
        self.n = 0

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.mg_grid_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            msg = self._L.mg_grid_last_error(self._h)
            msg = msg.decode() if msg else f"{what} failed ({rc})"
            raise (ValueError if rc == -4 else RuntimeError)(f"{what}: {msg}")

    def build(self, points, cell_size: float):
        """UniformGrid::build (grid.rs:101-153) over ``points`` ((n, 2) float32), copied to the device."""
        pts = _points(points)
        self.points = DeviceArray.from_numpy(pts)
        self.n = pts.shape[0]
        self.cell_size = float(cell_size)
        self._check(self._L.mg_grid_build(self._h, self.points.ptr, self.n, self.cell_size, None), "mg_grid_build")
        return self

    def meta(self) -> dict:
        m = np.zeros(8, dtype=np.int64)
        self._check(self._L.mg_grid_meta(self._h, m.ctypes.data_as(C.POINTER(C.c_int64))), "mg_grid_meta")
        return {"dense": bool(m[0]), "min_cx": int(m[1]), "min_cy": int(m[2]), "width": int(m[3]),
                "height": int(m[4]), "n_cell_start": int(m[5]), "n_indices": int(m[6]), "n_points": int(m[7])}

    def layout(self) -> dict:
        """Host copies: meta plus cell_start, indices and (Sparse) the (cx, cy) keys."""
        md = self.meta()
        cs = np.zeros(md["n_cell_start"], dtype=np.uint32)
        idx = np.zeros(max(md["n_indices"], 1), dtype=np.uint32)
        nk = 0 if md["dense"] else md["n_cell_start"] - 1
        keys = np.zeros((max(nk, 1), 2), dtype=np.int32)
        u32p = C.POINTER(C.c_uint32)
        self._check(self._L.mg_grid_copy(self._h, cs.ctypes.data_as(u32p), cs.size, idx.ctypes.data_as(u32p), idx.size,
                                         keys.ctypes.data_as(C.POINTER(C.c_int32)), nk), "mg_grid_copy")
        md.update(cell_start=cs, indices=idx[: md["n_indices"]], keys=keys[:nk])
        return md

    def _two_sizes(self, first_cap, call):
        """Runs call(out, cap) with a device output of cap u32 words, once more when it did not fit."""
        cap = max(int(first_cap), 1)
        for _ in range(2):
            out = DeviceArray(4 * cap)
            total = call(out, cap)
            if total <= cap:
                return out, total
            cap = total
        raise RuntimeError("output size changed between passes")

    def candidate_pairs(self):
        """for_each_candidate_pair (grid.rs:291-362): (k, 2) pairs (i < j) in emission order."""
        tot = C.c_uint64()

        def call(out, cap):
            self._check(self._L.mg_grid_candidate_pairs(self._h, out.ptr, cap // 2, C.byref(tot), None),
                        "mg_grid_candidate_pairs")
            return 2 * int(tot.value)
        out, total = self._two_sizes(16 * max(self.n, 1), call)
        return out.to_numpy(np.uint32, total).reshape(-1, 2)

    def neighbor_lists(self, h: float):
        """build_neighbor_lists (pbf.rs:262-288) over the points this grid was built from:
        (offsets (n + 1), neighbours)."""
        off = DeviceArray(4 * (self.n + 1))
        tot = C.c_uint64()

        def call(out, cap):
            self._check(self._L.mg_grid_neighbor_lists(self._h, self.points.ptr, float(h), off.ptr, out.ptr, cap,
                                                       C.byref(tot), None), "mg_grid_neighbor_lists")
            return int(tot.value)
        out, total = self._two_sizes(32 * max(self.n, 1), call)
        return off.to_numpy(np.uint32, self.n + 1), out.to_numpy(np.uint32, total)

    def near(self, queries):
        """for_each_near (grid.rs:371-399) of each query point: (offsets, candidates)."""
        q = _points(queries)
        dq = DeviceArray.from_numpy(q)
        off = DeviceArray(4 * (q.shape[0] + 1))
        tot = C.c_uint64()

        def call(out, cap):
            self._check(self._L.mg_grid_near(self._h, dq.ptr, q.shape[0], off.ptr, out.ptr, cap, C.byref(tot), None),
                        "mg_grid_near")
            return int(tot.value)
        out, total = self._two_sizes(32 * max(q.shape[0], 1), call)
        return off.to_numpy(np.uint32, q.shape[0] + 1), out.to_numpy(np.uint32, total)
