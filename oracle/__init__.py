"""TEST INFRASTRUCTURE — ctypes wrapper of the CPU oracle (oracle/liboracle_hns.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module;
the product path never does. The oracle restates the reference's algorithm on the host (see the
header of hns_oracle.cpp for the file:line it follows); `dtype=np.float64` runs the same
sequence of operations in double precision (the FP64 host oracle).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Dict, Optional

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_DIR, "liboracle_hns.so")
_lib = None


class OracleParams(C.Structure):
    _fields_ = [("expansionRate", C.c_float), ("temperatureRelease", C.c_float), ("buoyancyStrength", C.c_float),
                ("ambientTemp", C.c_float), ("vorticityScale", C.c_float), ("factorScale", C.c_float),
                ("disturbanceEnable", C.c_float), ("disturbanceStrength", C.c_float),
                ("disturbanceSwirl", C.c_float), ("disturbanceThreshold", C.c_float),
                ("disturbanceGain", C.c_float), ("disturbanceFrequency", C.c_float), ("substeps", C.c_int32),
                ("gravity", C.c_float), ("maskDensityMin", C.c_float), ("maskDensityMax", C.c_float),
                ("downloadVel", C.c_int32), ("simReset", C.c_int32)]


def build():
    subprocess.run(["make", "-s", "-C", _DIR], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        _lib.oracle_topo_create.restype = vp
        _lib.oracle_topo_create.argtypes = [vp, C.c_int64]
        _lib.oracle_topo_destroy.argtypes = [vp]
        _lib.oracle_offset.restype = C.c_int64
        _lib.oracle_offset.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32]
        _lib.oracle_leaf_order.argtypes = [vp, C.c_int64, vp]
        _lib.oracle_sample_nearest_f32.restype = C.c_float
        _lib.oracle_sample_nearest_f32.argtypes = [vp, vp, C.c_int32, C.c_int32, C.c_int32]
        _lib.oracle_sample_trilinear_f32.restype = C.c_float
        _lib.oracle_sample_trilinear_f32.argtypes = [vp, vp, C.c_float, C.c_float, C.c_float]
        _lib.oracle_sample_trilinear_vec_f32.argtypes = [vp, vp, C.c_float, C.c_float, C.c_float, vp]
        _lib.oracle_num_threads.restype = C.c_int
        _lib.oracle_set_num_threads.argtypes = [C.c_int]
        for suf in ("f32", "f64"):
            f = lambda n: getattr(_lib, "oracle_%s_%s" % (n, suf))
            f("advect_vector").argtypes = [vp, vp, vp, C.c_float, C.c_float, vp]
            f("advect_scalars").argtypes = [vp, vp, C.c_int, vp, vp, C.c_float, C.c_float, vp]
            f("advect_scalar").argtypes = [vp, vp, vp, vp, C.c_float, C.c_float, vp]
            f("vorticity").argtypes = [vp, vp, vp, C.c_float, C.c_float, C.c_float, C.c_float]
            f("divergence").argtypes = [vp, vp, vp, C.c_float]
            f("rbgs").argtypes = [vp, vp, vp, C.c_float, C.c_int, C.c_float]
            f("residual").argtypes = [vp, vp, vp, C.c_float, vp, vp]
            f("cg").restype = C.c_int
            f("cg").argtypes = [vp, vp, C.c_float, C.c_double, C.c_int, vp, vp]
            f("subtract_gradient").argtypes = [vp, vp, vp, vp, C.c_float, vp]
            f("enforce_collision").argtypes = [vp, vp, vp, C.c_float]
            f("combustion").argtypes = [C.c_int64, vp, vp, vp, vp, vp, C.c_float, C.c_float]
            f("buoyancy").argtypes = [C.c_int64, vp, vp, C.c_float, C.c_float, C.c_float]
            f("advect_scalars_adj").argtypes = [vp, vp, vp, vp, vp, vp, C.c_float, C.c_float]
            f("advect_vector_adj").argtypes = [vp, vp, vp, vp, C.c_float, C.c_float]
            f("divergence_adj").argtypes = [vp, vp, vp, C.c_float]
            f("gradient_adj").argtypes = [vp, vp, vp, C.c_float]
            f("rbgs_adj").argtypes = [vp, vp, vp, C.c_float, C.c_int, C.c_float]
            f("diff_forward").argtypes = [vp, vp, vp, C.c_int, C.c_float, C.c_float, C.c_int, C.c_float]
            f("diff_loss_grad").restype = C.c_double
            f("diff_loss_grad").argtypes = [vp, vp, vp, vp, C.c_int, C.c_float, C.c_float, C.c_int, C.c_float, vp, vp]
            f("advance").restype = C.c_int
            f("advance").argtypes = [vp, C.c_float, C.c_int, C.c_float, C.POINTER(OracleParams), vp, C.c_int, vp, vp,
                                     vp, C.c_int, vp, vp, vp, vp]
        _lib.oracle_advance_mgcg_f32.restype = C.c_int
        _lib.oracle_advance_mgcg_f32.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.POINTER(OracleParams), vp,
                                                 C.c_int, vp, vp, vp, C.c_int, vp, vp, vp, vp, vp, C.c_int]
        _lib.oracle_mg_create.restype = vp
        _lib.oracle_mg_create.argtypes = [vp, C.c_int64]
        _lib.oracle_mg_destroy.argtypes = [vp]
        _lib.oracle_mg_two_level.restype = C.c_int
        _lib.oracle_mg_two_level.argtypes = [vp]
        _lib.oracle_mg_blocks.restype = C.c_int64
        _lib.oracle_mg_blocks.argtypes = [vp]
        _lib.oracle_mg_vcycle.argtypes = [vp, vp, vp]
        _lib.oracle_mg_apply_level.argtypes = [vp, C.c_int, vp, C.c_float, vp]
        _lib.oracle_mg_apply_level4.argtypes = [vp, vp, vp]
        _lib.oracle_mg_blocks_of.argtypes = [vp, vp, vp]
        _lib.oracle_mgcg.restype = C.c_int
        _lib.oracle_mgcg.argtypes = [vp, vp, C.c_float, C.c_double, C.c_int, vp, vp]
    return _lib


def set_threads(n: int):
    load().oracle_set_num_threads(int(n))


def num_threads() -> int:
    return int(load().oracle_num_threads())


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


def leaf_order(origins: np.ndarray) -> np.ndarray:
    o = np.ascontiguousarray(origins, dtype=np.int32).reshape(-1, 3)
    perm = np.empty(len(o), np.int64)
    load().oracle_leaf_order(o.ctypes.data, len(o), perm.ctypes.data)
    return perm


class Oracle:
    """Topology + kernels of the reference step on the host. Arrays are leaf-dense; Vec3 fields
    are (N, 3) AoS as the reference stores them."""

    def __init__(self, coords: np.ndarray, dtype=np.float32):
        self.coords = np.ascontiguousarray(coords, dtype=np.int32).reshape(-1, 3)
        self.n = len(self.coords)
        self.dtype = np.dtype(dtype)
        self.suf = "f32" if self.dtype == np.float32 else "f64"
        self._t = load().oracle_topo_create(self.coords.ctypes.data, self.n)
        if not self._t:
            raise ValueError("oracle: coords are not a leaf-dense payload")

    def __del__(self):
        if getattr(self, "_t", None):
            load().oracle_topo_destroy(self._t)
            self._t = None

    def _f(self, name):
        return getattr(load(), "oracle_%s_%s" % (name, self.suf))

    def _a(self, a, shape=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if shape is not None:
            a = a.reshape(shape)
        return a

    def offset(self, x, y, z) -> int:
        return int(load().oracle_offset(self._t, int(x), int(y), int(z)))

    def nearest(self, data, x, y, z) -> float:
        d = np.ascontiguousarray(data, np.float32)
        return float(load().oracle_sample_nearest_f32(self._t, d.ctypes.data, int(x), int(y), int(z)))

    def trilinear(self, data, x, y, z) -> float:
        d = np.ascontiguousarray(data, np.float32)
        return float(load().oracle_sample_trilinear_f32(self._t, d.ctypes.data, x, y, z))

    def trilinear_vec(self, data, x, y, z) -> np.ndarray:
        d = np.ascontiguousarray(data, np.float32).reshape(-1, 3)
        out = np.empty(3, np.float32)
        load().oracle_sample_trilinear_vec_f32(self._t, d.ctypes.data, x, y, z, out.ctypes.data)
        return out

    def advect_vector(self, vel, dt, inv_dx, sdf=None):
        v = self._a(vel, (self.n, 3))
        out = np.empty_like(v)
        s = self._a(sdf) if sdf is not None else None
        self._f("advect_vector")(self._t, v.ctypes.data, out.ctypes.data, dt, inv_dx, _p(s))
        return out

    def advect_scalars(self, vel, fields, dt, inv_dx, sdf=None):
        v = self._a(vel, (self.n, 3))
        ins = [self._a(f) for f in fields]
        outs = [np.empty_like(f) for f in ins]
        pin = (C.c_void_p * max(1, len(ins)))(*[f.ctypes.data for f in ins])
        pout = (C.c_void_p * max(1, len(outs)))(*[f.ctypes.data for f in outs])
        s = self._a(sdf) if sdf is not None else None
        self._f("advect_scalars")(self._t, v.ctypes.data, len(ins), C.addressof(pin), C.addressof(pout), dt, inv_dx, _p(s))
        return outs

    def advect_scalar(self, vel, field, dt, inv_dx, sdf=None):
        v = self._a(vel, (self.n, 3))
        f = self._a(field)
        out = np.empty_like(f)
        s = self._a(sdf) if sdf is not None else None
        self._f("advect_scalar")(self._t, v.ctypes.data, f.ctypes.data, out.ctypes.data, dt, inv_dx, _p(s))
        return out

    def vorticity(self, vel, dt, inv_dx, scale, factor_scale):
        v = self._a(vel, (self.n, 3))
        out = np.empty_like(v)
        self._f("vorticity")(self._t, v.ctypes.data, out.ctypes.data, dt, inv_dx, scale, factor_scale)
        return out

    def divergence(self, vel, inv_dx):
        v = self._a(vel, (self.n, 3))
        out = np.empty(self.n, self.dtype)
        self._f("divergence")(self._t, v.ctypes.data, out.ctypes.data, inv_dx)
        return out

    def rbgs(self, div, p, dx, colour, omega):
        """One colour sweep, in place on p (returned)."""
        d = self._a(div)
        q = self._a(p).copy()
        self._f("rbgs")(self._t, d.ctypes.data, q.ctypes.data, dx, int(colour), omega)
        return q

    def residual(self, p, div, dx):
        r2, d2 = C.c_double(), C.c_double()
        pa, da = self._a(p), self._a(div)  # keep the arrays alive across the call
        self._f("residual")(self._t, pa.ctypes.data, da.ctypes.data, dx, C.byref(r2), C.byref(d2))
        return r2.value, d2.value

    def cg(self, div, dx, tol=1e-5, max_iters=100000):
        """CG solve of L p = div from p = 0; returns (p, iterations, recurrence relative residual)."""
        d = self._a(div)
        p = np.empty(self.n, self.dtype)
        rel = C.c_double()
        it = self._f("cg")(self._t, d.ctypes.data, dx, float(tol), int(max_iters), p.ctypes.data, C.byref(rel))
        return p, int(it), rel.value

    # ---- reverse mode of the differentiable step (FP64 host tape when dtype is float64) ----
    def advect_scalars_adj(self, vel, field, gout, dt, inv_dx):
        """Returns (d/d field, d/d vel) of sum(gout * advect_scalars(field; vel))."""
        v, f, g = self._a(vel, (self.n, 3)), self._a(field), self._a(gout)
        gin = np.zeros(self.n, self.dtype)
        gvel = np.zeros((self.n, 3), self.dtype)
        self._f("advect_scalars_adj")(self._t, v.ctypes.data, f.ctypes.data, g.ctypes.data, gin.ctypes.data,
                                      gvel.ctypes.data, dt, inv_dx)
        return gin, gvel

    def advect_vector_adj(self, vel, gout, dt, inv_dx):
        v, g = self._a(vel, (self.n, 3)), self._a(gout, (self.n, 3))
        gvel = np.zeros((self.n, 3), self.dtype)
        self._f("advect_vector_adj")(self._t, v.ctypes.data, g.ctypes.data, gvel.ctypes.data, dt, inv_dx)
        return gvel

    def divergence_adj(self, gdiv, inv_dx):
        g = self._a(gdiv)
        gvel = np.zeros((self.n, 3), self.dtype)
        self._f("divergence_adj")(self._t, g.ctypes.data, gvel.ctypes.data, inv_dx)
        return gvel

    def gradient_adj(self, gu2, inv_dx):
        """d/dp of sum(gu2 . u2) for u2 = u1 - grad p (minus the transpose of the gradient)."""
        g = self._a(gu2, (self.n, 3))
        gp = np.zeros(self.n, self.dtype)
        self._f("gradient_adj")(self._t, g.ctypes.data, gp.ctypes.data, inv_dx)
        return gp

    def rbgs_adj(self, gp, dx, K, omega):
        """d/d div of sum(gp * p_K(div)) for K RBGS iterations from p = 0."""
        g = self._a(gp)
        gdiv = np.zeros(self.n, self.dtype)
        self._f("rbgs_adj")(self._t, g.ctypes.data, gdiv.ctypes.data, dx, int(K), omega)
        return gdiv

    def diff_forward(self, vel, rho, steps, dt, dx, K, omega):
        v, r = self._a(vel, (self.n, 3)).copy(), self._a(rho).copy()
        self._f("diff_forward")(self._t, v.ctypes.data, r.ctypes.data, int(steps), dt, dx, int(K), omega)
        return v, r

    def diff_loss_grad(self, vel, rho, target, steps, dt, dx, K, omega):
        """L = sum (rho_S - target)^2 and (dL/dvel0, dL/drho0)."""
        v, r, tg = self._a(vel, (self.n, 3)), self._a(rho), self._a(target)
        gv = np.zeros((self.n, 3), self.dtype)
        gr = np.zeros(self.n, self.dtype)
        loss = self._f("diff_loss_grad")(self._t, v.ctypes.data, r.ctypes.data, tg.ctypes.data, int(steps), dt, dx,
                                         int(K), omega, gv.ctypes.data, gr.ctypes.data)
        return loss, gv, gr

    def subtract_gradient(self, vel, p, inv_dx, sdf=None):
        v = self._a(vel, (self.n, 3))
        out = np.empty_like(v)
        s = self._a(sdf) if sdf is not None else None
        pa = self._a(p)
        self._f("subtract_gradient")(self._t, v.ctypes.data, pa.ctypes.data, out.ctypes.data, inv_dx, _p(s))
        return out

    def enforce_collision(self, vel, sdf, voxel_size):
        v = self._a(vel, (self.n, 3)).copy()
        sa = self._a(sdf)
        self._f("enforce_collision")(self._t, v.ctypes.data, sa.ctypes.data, voxel_size)
        return v

    def combustion(self, fuel, waste, temp, div, flame, gain, expansion):
        arrs = [self._a(a).copy() for a in (fuel, waste, temp, div, flame)]
        self._f("combustion")(self.n, *[a.ctypes.data for a in arrs], gain, expansion)
        return arrs  # fuel, waste, temp, div, flame

    def buoyancy(self, vel, temp, dt, ambient, beta):
        v = self._a(vel, (self.n, 3)).copy()
        ta = self._a(temp)
        self._f("buoyancy")(self.n, v.ctypes.data, ta.ctypes.data, dt, ambient, beta)
        return v

    def advance(self, voxel_size, iterations, dt, params: Dict[str, float], velocity, fields: Dict[str, np.ndarray],
                sources: Optional[Dict[str, np.ndarray]] = None, src_vel=None, sdf=None, mgcg_tol: Optional[float] = None,
                solve_iterations: Optional[list] = None):
        """HNS_Advance on host copies. `fields` maps state names to arrays (the context's state at call
        start); the combustion fields it lacks are created as the reference's needOrMake does.
        Returns (velocity, fields) after the call. With `mgcg_tol` (FP32 only) every pressure solve is
        the multigrid-preconditioned CG to that relative residual instead of `iterations` RBGS
        iterations; their iteration counts are appended to `solve_iterations`."""
        P = OracleParams()
        for k, _ in OracleParams._fields_:
            v = params[k]
            setattr(P, k, int(v) if k in ("substeps", "downloadVel", "simReset") else float(v))
        vel = self._a(velocity, (self.n, 3)).copy()
        names = list(fields.keys())
        arrs = [self._a(fields[n]).copy() for n in names]
        present = [1] * len(names)
        for extra in ("fuel", "waste", "temperature", "flame"):
            if extra not in fields:
                names.append(extra)
                arrs.append(np.zeros(self.n, self.dtype))
                present.append(0)
        cn = (C.c_char_p * len(names))(*[n.encode() for n in names])
        cf = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        cp = np.asarray(present, np.int32)
        sources = sources or {}
        snames = list(sources.keys())
        sarrs = [self._a(sources[n]) for n in snames]
        csn = (C.c_char_p * max(1, len(snames)))(*[n.encode() for n in snames])
        csa = (C.c_void_p * max(1, len(sarrs)))(*[a.ctypes.data for a in sarrs])
        sv = self._a(src_vel, (self.n, 3)) if src_vel is not None else None
        sd = self._a(sdf) if sdf is not None else None
        if mgcg_tol is None:
            st = self._f("advance")(self._t, voxel_size, int(iterations), dt, C.byref(P), vel.ctypes.data, len(names),
                                    C.addressof(cn), C.addressof(cf), cp.ctypes.data, len(snames), C.addressof(csn),
                                    C.addressof(csa), _p(sv), _p(sd))
        else:
            if self.dtype != np.float32:
                raise ValueError("the multigrid-preconditioned step is FP32")
            its = np.zeros(64, np.int32)
            st = load().oracle_advance_mgcg_f32(self._t, voxel_size, float(mgcg_tol), dt, C.byref(P), vel.ctypes.data,
                                                len(names), C.addressof(cn), C.addressof(cf), cp.ctypes.data,
                                                len(snames), C.addressof(csn), C.addressof(csa), _p(sv), _p(sd),
                                                its.ctypes.data, len(its))
            if solve_iterations is not None:
                solve_iterations.extend(int(v) for v in its[its > 0])
        if st != 0:
            raise RuntimeError("oracle advance failed: %d" % st)
        return vel, dict(zip(names, arrs))


class Multigrid:
    """The V-cycle preconditioner of the library's multigrid-preconditioned solve and the flexible
    CG on it (hns_oracle_mg.cpp), FP32 on leaf-dense level-0 vectors of a payload whose leaves are
    given by their origins (one per 512 voxels, payload order)."""

    def __init__(self, coords: np.ndarray):
        c = np.ascontiguousarray(coords, dtype=np.int32).reshape(-1, 3)
        self.origins = np.ascontiguousarray(c[::512])
        self.L = len(self.origins)
        self.n = 512 * self.L
        self._g = load().oracle_mg_create(self.origins.ctypes.data, self.L)

    def __del__(self):
        if getattr(self, "_g", None):
            load().oracle_mg_destroy(self._g)
            self._g = None

    @property
    def two_level(self) -> bool:
        return bool(load().oracle_mg_two_level(self._g))

    @property
    def num_blocks(self) -> int:
        return int(load().oracle_mg_blocks(self._g))

    def vcycle(self, r) -> np.ndarray:
        r = np.ascontiguousarray(r, np.float32)
        z = np.empty(self.n, np.float32)
        load().oracle_mg_vcycle(self._g, r.ctypes.data, z.ctypes.data)
        return z

    def apply_level(self, level: int, u, scale: float) -> np.ndarray:
        """scale (6 u - face neighbours) on level `level` (0..3, (8 >> level)^3 cells per leaf)."""
        u = np.ascontiguousarray(u, np.float32)
        out = np.empty_like(u)
        load().oracle_mg_apply_level(self._g, int(level), u.ctypes.data, float(scale), out.ctypes.data)
        return out

    def apply_level4(self, e) -> np.ndarray:
        e = np.ascontiguousarray(e, np.float64)
        out = np.empty_like(e)
        load().oracle_mg_apply_level4(self._g, e.ctypes.data, out.ctypes.data)
        return out

    def blocks(self):
        """(level-4 block of each leaf, colour of each block)."""
        blk = np.empty(self.L, np.int32)
        col = np.empty(self.num_blocks, np.int32)
        load().oracle_mg_blocks_of(self._g, blk.ctypes.data, col.ctypes.data)
        return blk, col

    def solve(self, div, dx: float, tol: float = 1e-5, max_iters: int = 1000):
        """Flexible CG on the V-cycle for A p = -div dx^2 from p = 0: (p, iterations, rel residual)."""
        d = np.ascontiguousarray(div, np.float32)
        p = np.empty(self.n, np.float32)
        rel = C.c_double()
        it = load().oracle_mgcg(self._g, d.ctypes.data, float(dx), float(tol), int(max_iters), p.ctypes.data,
                                C.byref(rel))
        return p, int(it), float(rel.value)
