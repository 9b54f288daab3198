"""hnanosolver_amd — MI355X (gfx950) sparse smoke/fire step behind HNanoSolver's launch API.

Python mirror of the reference's host interface over the C ABI in include/hns.h:

    data = GridIndexedData(coords)                    # HNS::GridIndexedData (GridData.hpp:16)
    data.add_value_block("velocity", vel)             # addValueBlock<T> (GridData.hpp:61)
    topo = create_index_grid(data, voxel_size)        # CreateIndexGrid (HNanoSolver.cu:507)
    ctx = ensure_context(None, data, topo, voxel_size, has_collision)   # HNS_EnsureContext (:58)
    upload_sources(ctx, data)                         # HNS_UploadSources (:158)
    advance(ctx, iterations, dt, params)              # HNS_Advance (:223)
    download_for_output(ctx, data)                    # HNS_DownloadForOutput (:471)
    mask = get_active_mask(ctx)                       # HNS_GetActiveMask (:459)
    destroy_context(ctx)                              # HNS_DestroyContext (:51)

Failures raise HNSError (the reference throws std::runtime_error). There is no CPU
fallback: every call runs the HIP kernels of libhns.so or raises.
"""
from __future__ import annotations

import ctypes as C
from collections import OrderedDict
from typing import Dict, Optional

import numpy as np

from . import _lib
from ._lib import hns_params, hns_solve_options, hns_solve_stats

HNS_OK = 0
STATUS = {0: "HNS_OK", -1: "HNS_ERR_INVALID_ARGUMENT", -2: "HNS_ERR_EMPTY", -3: "HNS_ERR_NULL_CONTEXT",
          -4: "HNS_ERR_TOPOLOGY", -5: "HNS_ERR_LEAF_ORDER", -6: "HNS_ERR_NO_VELOCITY", -7: "HNS_ERR_HIP",
          -8: "HNS_ERR_OUT_OF_MEMORY", -9: "HNS_ERR_NOT_FOUND", -10: "HNS_ERR_NO_DEVICE",
          -11: "HNS_ERR_NOT_CONVERGED", -12: "HNS_ERR_NON_FINITE"}
HNS_FLOAT, HNS_VEC3F = 1, 3
SOLVERS = {"rbgs": 0, "rbgs_tol": 1, "cg": 2, "pcg_learned": 3, "mgcg": 4}
KERNEL_CLASSES = ["advect_velocity", "vorticity_forces", "divergence_combustion", "rbgs", "residual", "gradient",
                  "advect_scalars", "cg_product", "cg_update", "precond", "other"]


class HNSError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__("%s (%d): %s" % (STATUS.get(status, "?"), status, message))
        self.status = status


def lib():
    return _lib.load()


def _check(st: int):
    if st != HNS_OK:
        raise HNSError(st, lib().hns_last_error().decode(errors="replace"))


# --------------------------------------------------------------------------------------
# payload
# --------------------------------------------------------------------------------------
class GridIndexedData:
    """Leaf-dense payload: coords (N, 3) int32 plus named float / Vec3f blocks in insertion order."""

    def __init__(self, coords: np.ndarray):
        self.coords = np.ascontiguousarray(coords, dtype=np.int32).reshape(-1, 3)
        self.blocks: "OrderedDict[str, np.ndarray]" = OrderedDict()

    def size(self) -> int:
        return int(self.coords.shape[0])

    def add_value_block(self, name: str, values: Optional[np.ndarray] = None, vec3: Optional[bool] = None):
        if name in self.blocks:
            raise ValueError("Block with name '%s' already exists!" % name)
        if values is None:
            values = np.zeros((self.size(), 3) if vec3 else self.size(), np.float32)
        a = np.ascontiguousarray(values, dtype=np.float32)
        if a.ndim == 2 and a.shape[1] == 3:
            pass
        elif a.ndim == 1:
            pass
        else:
            raise ValueError("block '%s' must be (N,) or (N, 3)" % name)
        if a.shape[0] != self.size():
            raise ValueError("block '%s' has %d elements, payload has %d" % (name, a.shape[0], self.size()))
        self.blocks[name] = a
        return a

    def get_blocks_of_type(self, vec3: bool):
        return [n for n, a in self.blocks.items() if (a.ndim == 2) == vec3]

    def __getitem__(self, name):
        return self.blocks[name]

    def _c(self):
        """C view (hns_grid_data) valid while self is alive."""
        names = [n.encode() for n in self.blocks]
        arr = (_lib.hns_value_block * max(1, len(self.blocks)))()
        for i, (n, a) in enumerate(self.blocks.items()):
            arr[i].name = names[i]
            arr[i].type = HNS_VEC3F if a.ndim == 2 else HNS_FLOAT
            arr[i].values = a.ctypes.data
        g = _lib.hns_grid_data()
        g.size = self.size()
        g.coords = self.coords.ctypes.data_as(_lib.c_int32_p)
        g.num_blocks = len(self.blocks)
        g.blocks = arr
        g._keep = (names, arr)
        return g


def make_params(**kw) -> hns_params:
    p = hns_params()
    from .scenes import default_params
    d = default_params(**kw)
    for k, v in d.items():
        setattr(p, k, int(v) if k in ("substeps", "downloadVel", "simReset") else float(v))
    return p


# --------------------------------------------------------------------------------------
# topology
# --------------------------------------------------------------------------------------
class Topology:
    """Device-resident leaf topology (replaces nanovdb::GridHandle<cuda::DeviceBuffer>)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    @property
    def handle(self):
        return self._h

    def info(self):
        n = C.c_size_t()
        l = C.c_size_t()
        _check(lib().hns_topology_info(self._h, C.byref(n), C.byref(l)))
        return n.value, l.value

    def neighbours(self) -> np.ndarray:
        _, L = self.info()
        out = np.empty((L, 27), np.int32)
        _check(lib().hns_topology_neighbours(self._h, out.ctypes.data_as(_lib.c_int32_p), out.size))
        return out

    def builds(self):
        """(device builds, reuses): calls of create_index_grid that kept this handle count as reuses."""
        b, r = C.c_int64(), C.c_int64()
        _check(lib().hns_topology_builds(self._h, C.byref(b), C.byref(r)))
        return b.value, r.value

    def destroy(self):
        if self._h:
            h = C.c_void_p(self._h.value if isinstance(self._h, C.c_void_p) else self._h)
            _check(lib().hns_topology_destroy(C.byref(h)))
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def create_index_grid(data: GridIndexedData, voxel_size: float, topo: Optional[Topology] = None) -> Topology:
    """CreateIndexGrid. With `topo` (the previous cook's handle) the handle is kept when the payload's
    leaf origins and voxel size are unchanged, and rebuilt in place otherwise."""
    g = data._c()
    if topo is None or not topo.handle:
        h = C.c_void_p()
        _check(lib().hns_create_index_grid(C.byref(g), float(voxel_size), C.byref(h)))
        return Topology(h)
    h = topo.handle if isinstance(topo.handle, C.c_void_p) else C.c_void_p(topo.handle)
    _check(lib().hns_create_index_grid(C.byref(g), float(voxel_size), C.byref(h)))
    topo._h = h
    return topo


def leaf_order(origins: np.ndarray) -> np.ndarray:
    o = np.ascontiguousarray(origins, dtype=np.int32).reshape(-1, 3)
    perm = np.empty(len(o), np.uint64)
    _check(lib().hns_leaf_order(o.ctypes.data_as(_lib.c_int32_p), len(o), perm.ctypes.data_as(C.POINTER(C.c_uint64))))
    return perm.astype(np.int64)


# --------------------------------------------------------------------------------------
# context
# --------------------------------------------------------------------------------------
class Context:
    """Opaque SimContext handle (src/Cuda/HNanoSolver.cu:27-48)."""

    def __init__(self):
        self._h = C.c_void_p()
        self._topo = None

    @property
    def handle(self):
        return self._h

    def __bool__(self):
        return bool(self._h.value)

    def info(self):
        n, l, s = C.c_size_t(), C.c_size_t(), C.c_int32()
        _check(lib().hns_context_info(self._h, C.byref(n), C.byref(l), C.byref(s)))
        return n.value, l.value, s.value

    def has_field(self, name: str) -> bool:
        return bool(lib().hns_has_field(self._h, name.encode()))

    def set_stream(self, stream_ptr: int):
        _check(lib().hns_set_stream(self._h, C.c_void_p(stream_ptr)))

    def set_preconditioner(self, pc):
        """HNS_SOLVER_PCG_LEARNED solves of this context use `pc` (a precond.Preconditioner, kept
        alive by the context; None detaches)."""
        _check(lib().hns_set_preconditioner(self.handle, pc.handle if pc is not None else None))
        self._pc = pc

    def set_timing(self, on: bool):
        _check(lib().hns_set_timing(self._h, 1 if on else 0))

    def reset_timing(self):
        _check(lib().hns_reset_timing(self._h))

    def timing(self) -> Dict[str, tuple]:
        out = {}
        for k, name in enumerate(KERNEL_CLASSES):
            ms, n = C.c_double(), C.c_int64()
            _check(lib().hns_get_timing(self._h, k, C.byref(ms), C.byref(n)))
            out[name] = (ms.value, n.value)
        return out

    def __del__(self):
        try:
            destroy_context(self)
        except Exception:
            pass


def ensure_context(ctx: Optional[Context], data: GridIndexedData, topo: Topology, voxel_size: float,
                   has_collision: bool = False) -> Context:
    if ctx is None:
        ctx = Context()
    g = data._c()
    _check(lib().hns_ensure_context(C.byref(ctx._h), C.byref(g), topo.handle, float(voxel_size),
                                    1 if has_collision else 0))
    ctx._topo = topo
    return ctx


def upload_sources(ctx: Context, data: GridIndexedData):
    g = data._c()
    _check(lib().hns_upload_sources(ctx.handle, C.byref(g)))


def advance(ctx: Context, iterations: int, dt: float, params: hns_params, solver: Optional[str] = None,
            tol: float = 1e-5, max_iters: int = 100000, check_every: int = 0, stats: bool = False):
    """HNS_Advance; with `solver` or `stats`, returns hns_solve_stats of the last pressure solve."""
    if solver is None and not stats:
        _check(lib().hns_advance(ctx.handle, int(iterations), float(dt), C.byref(params)))
        return None
    o = hns_solve_options(C.sizeof(hns_solve_options), SOLVERS[solver or "rbgs"], float(tol), int(max_iters),
                          int(check_every))
    s = hns_solve_stats(C.sizeof(hns_solve_stats), 0, -1.0, 0.0)
    st = lib().hns_advance_ex(ctx.handle, int(iterations), float(dt), C.byref(params), C.byref(o), C.byref(s))
    if st != -11:
        _check(st)
    return s


def download_for_output(ctx: Context, data: GridIndexedData):
    g = data._c()
    _check(lib().hns_download_for_output(ctx.handle, C.byref(g)))


def get_active_mask(ctx: Optional[Context]) -> np.ndarray:
    h = ctx.handle if ctx is not None else None
    n = C.c_size_t()
    _check(lib().hns_get_active_mask(h, None, 0, C.byref(n)))
    out = np.empty(n.value, np.uint8)
    if n.value:
        _check(lib().hns_get_active_mask(h, out.ctypes.data_as(C.POINTER(C.c_uint8)), n.value, C.byref(n)))
    return out


def destroy_context(ctx: Optional[Context]):
    if ctx is None or not ctx._h:
        return
    _check(lib().hns_destroy_context(C.byref(ctx._h)))


# --------------------------------------------------------------------------------------
# standalone entry points (HNanoAdvect, HNanoAdvectVelocity, HNanoProjectNonDivergent SOPs)
# --------------------------------------------------------------------------------------
def advect_index_grid(data: GridIndexedData, dt: float, voxel_size: float):
    g = data._c()
    _check(lib().hns_advect_index_grid(C.byref(g), float(dt), float(voxel_size), None))


def advect_index_grid_velocity(data: GridIndexedData, dt: float, voxel_size: float):
    g = data._c()
    _check(lib().hns_advect_index_grid_velocity(C.byref(g), float(dt), float(voxel_size), None))


def divergence(data: GridIndexedData, voxel_size: float):
    g = data._c()
    _check(lib().hns_divergence(C.byref(g), float(voxel_size), None))


def project_non_divergent(data: GridIndexedData, iterations: int, voxel_size: float):
    g = data._c()
    _check(lib().hns_project_non_divergent(C.byref(g), int(iterations), float(voxel_size), None))
