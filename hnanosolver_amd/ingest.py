"""Host ingestion and output scatter (hns_payload_build / hns_scatter_leaves in include/hns.h).

The SOP's per-cook host work (SOP_HNanoSolver.cpp:276-301 and :345-388) over leaf grids, the form
OpenVDB and NanoVDB keep grids in:

    vel = LeafGrid.from_voxels("vel", coords, values)              # or LeafGrid(name, origins, values, active)
    payload = build_payload([density, vel, sdf, src], domain=1, padding=2)
    data = payload.data                                            # GridIndexedData over the payload memory
    topo = create_index_grid(data, voxel_size, topo)               # kept when the leaves did not change
    ...
    out = scatter_leaves(data, "density", background=0.0)          # LeafGrid of the set-on voxels
    grid = write_leaf_grid(data, "density", payload.domain_mask()) # writeIndexGrid: domain topology

The payload is built in libhns (C++); it lives in pinned host memory when a device is visible, so the
context's uploads of it are DMA transfers.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import HNS_FLOAT, HNS_VEC3F, GridIndexedData, _check, _lib, lib
from .topology import LEAF_LOCAL

ROLE_VALUE, ROLE_SDF = 0, 1


class LeafGrid:
    """A grid as its leaves: origins (L, 3) int32 (multiples of 8), values (L, 512) or (L, 512, 3)
    float32 in offset order x<<6 | y<<3 | z, active (L, 8) uint64 bit masks (None = all active)."""

    def __init__(self, name: str, origins, values, active=None, role: int = ROLE_VALUE, background: float = 0.0):
        self.name = name
        self.origins = np.ascontiguousarray(origins, np.int32).reshape(-1, 3)
        L = len(self.origins)
        v = np.ascontiguousarray(values, np.float32)
        if v.shape not in ((L, 512), (L, 512, 3)):
            raise ValueError("values of '%s' must be (L, 512) or (L, 512, 3), got %s" % (name, v.shape))
        self.values = v
        self.active = None if active is None else np.ascontiguousarray(active, np.uint64).reshape(L, 8)
        self.role = role
        self.background = float(background)

    @property
    def vec3(self) -> bool:
        return self.values.ndim == 3

    @classmethod
    def from_voxels(cls, name, coords, values, background: float = 0.0, role: int = ROLE_VALUE):
        """Leaves touched by the given voxels; those voxels active with their values, every other
        voxel of those leaves inactive at the background (openvdb setValueOn on an empty grid)."""
        coords = np.asarray(coords, np.int64).reshape(-1, 3)
        values = np.asarray(values, np.float32)
        vec3 = values.ndim == 2
        org = (coords >> 3) << 3
        uniq, inv = np.unique(org, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        L = len(uniq)
        off = ((coords[:, 0] & 7) << 6) | ((coords[:, 1] & 7) << 3) | (coords[:, 2] & 7)
        vals = np.full((L, 512, 3) if vec3 else (L, 512), background, np.float32)
        vals[inv, off] = values
        act = np.zeros((L, 8), np.uint64)
        np.bitwise_or.at(act, (inv, off >> 6), np.left_shift(np.uint64(1), (off & 63).astype(np.uint64)))
        return cls(name, uniq.astype(np.int32), vals, act, role, background)

    def voxels(self):
        """(coords (M, 3), values) of the active voxels."""
        L = len(self.origins)
        c = (self.origins[:, None, :] + LEAF_LOCAL[None]).reshape(-1, 3)
        v = self.values.reshape((L * 512, 3) if self.vec3 else (L * 512,))
        if self.active is None:
            return c, v
        bits = ((self.active[:, np.arange(512) >> 6] >> (np.arange(512) & 63).astype(np.uint64)) & np.uint64(1))
        m = bits.reshape(-1).astype(bool)
        return c[m], v[m]

    def _c(self):
        g = _lib.hns_leaf_grid()
        g.name = self.name.encode()
        g.type = HNS_VEC3F if self.vec3 else HNS_FLOAT
        g.role = self.role
        g.num_leaves = len(self.origins)
        g.origins = self.origins.ctypes.data_as(_lib.c_int32_p)
        g.values = self.values.ctypes.data_as(_lib.c_float_p)
        g.active = None if self.active is None else self.active.ctypes.data_as(C.POINTER(C.c_uint64))
        return g


class Payload:
    """hns_payload: the leaf-dense payload of one cook. `data` is a GridIndexedData whose arrays view
    the payload's memory (valid while this object lives)."""

    def __init__(self, handle):
        self._h = handle
        v = _lib.hns_grid_data()
        pinned = C.c_int32()
        _check(lib().hns_payload_view(self._h, C.byref(v), C.byref(pinned)))
        self.pinned = bool(pinned.value)
        n = int(v.size)
        coords = np.ctypeslib.as_array(v.coords, shape=(n, 3))
        d = GridIndexedData(coords)
        for i in range(v.num_blocks):
            b = v.blocks[i]
            shape = (n, 3) if b.type == HNS_VEC3F else (n,)
            arr = np.ctypeslib.as_array(C.cast(b.values, _lib.c_float_p), shape=shape)
            d.add_value_block(b.name.decode(), arr)
        self.data = d

    def domain_mask(self) -> np.ndarray:
        """hns_payload_domain_mask: the domain's active voxels, (L, 8) uint64 in payload leaf order
        (the dilated domain grid united with the SDF's active voxels)."""
        L = len(self.data.coords) // 512
        m = np.zeros((L, 8), np.uint64)
        _check(lib().hns_payload_domain_mask(self._h, m.ctypes.data_as(C.POINTER(C.c_uint64)), L))
        return m

    def close(self):
        if self._h:
            _check(lib().hns_payload_destroy(C.byref(self._h)))
            self._h = C.c_void_p()
            self.data = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_payload(grids: Sequence[LeafGrid], domain: int, padding: int) -> Payload:
    """hns_payload_build: domain = grids[domain]'s active voxels dilated by `padding` (26-neighbourhood)
    plus the first SDF grid's active voxels; one block per grid name (the first grid of a name wins)."""
    arr = (_lib.hns_leaf_grid * len(grids))(*[g._c() for g in grids])
    h = C.c_void_p()
    _check(lib().hns_payload_build(arr, len(grids), int(domain), int(padding), C.byref(h)))
    return Payload(h)


def scatter_leaves(data: GridIndexedData, name: str, mask: Optional[np.ndarray] = None,
                   background: float = 0.0) -> LeafGrid:
    """hns_scatter_leaves: the payload field `name` with the masked voxels set on (all voxels when
    mask is None, as HNS_GetActiveMask's mask is identically 1), as a LeafGrid."""
    g = data._c()
    m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    mp = None if m is None else m.ctypes.data_as(C.POINTER(C.c_uint8))
    n = C.c_size_t()
    _check(lib().hns_scatter_leaves(C.byref(g), name.encode(), mp, float(background), None, None, None, 0,
                                    C.byref(n)))
    L = n.value
    vec3 = data[name].ndim == 2
    origins = np.zeros((L, 3), np.int32)
    values = np.zeros((L, 512, 3) if vec3 else (L, 512), np.float32)
    active = np.zeros((L, 8), np.uint64)
    _check(lib().hns_scatter_leaves(C.byref(g), name.encode(), mp, float(background),
                                    origins.ctypes.data_as(_lib.c_int32_p), values.ctypes.data_as(_lib.c_float_p),
                                    active.ctypes.data_as(C.POINTER(C.c_uint64)), L, C.byref(n)))
    return LeafGrid(name, origins, values, active, background=background)


def write_leaf_grid(data: GridIndexedData, name: str, domain_active: Optional[np.ndarray] = None) -> LeafGrid:
    """hns_write_leaf_grid (IndexGridBuilder::writeIndexGrid): the payload field `name` as a LeafGrid
    of every payload leaf, its buffers the payload's values (active or not) and its active voxels
    domain_active's ((L, 8) uint64, as Payload.domain_mask gives them; None = every voxel)."""
    g = data._c()
    L = len(data.coords) // 512
    vec3 = data[name].ndim == 2
    origins = np.zeros((L, 3), np.int32)
    values = np.zeros((L, 512, 3) if vec3 else (L, 512), np.float32)
    active = np.zeros((L, 8), np.uint64)
    dm = None
    if domain_active is not None:
        dm = np.ascontiguousarray(domain_active, np.uint64).reshape(L, 8)
    _check(lib().hns_write_leaf_grid(C.byref(g), name.encode(),
                                     None if dm is None else dm.ctypes.data_as(C.POINTER(C.c_uint64)),
                                     origins.ctypes.data_as(_lib.c_int32_p), values.ctypes.data_as(_lib.c_float_p),
                                     active.ctypes.data_as(C.POINTER(C.c_uint64))))
    return LeafGrid(name, origins, values, active)
