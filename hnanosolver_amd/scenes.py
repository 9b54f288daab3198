"""Seeded synthetic scenes for the sparse smoke/fire step.

Every value is a pure function of (seed, global voxel coordinate), so results do not
depend on leaf order or on the device. The random source is SplitMix64 (Steele, Lea and
Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014) applied to packed
coordinates; curl noise is the curl of a vector potential made of octaves of smooth value
noise (smoothstep-weighted lattice interpolation), evaluated on a dense grid over the
bounding box and sampled at the leaf-dense voxels.

Domains are leaf-dense, as the reference builds them (src/Utils/GridBuilder.hpp:164-174,
:241): every voxel of every 8^3 leaf touched by the shape is simulated, in the leaf order
NanoVDB's voxelsToGrid produces (nanovdb/tools/cuda/PointsToGrid.cuh:596-651).
"""
from __future__ import annotations

import dataclasses
from typing import Dict, Optional

import numpy as np

from .topology import coords_from_leaves, leaf_order

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser over a uint64 array (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = x.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def _pack(x, y, z) -> np.ndarray:
    m = np.uint64((1 << 21) - 1)
    x = np.asarray(x).astype(np.int64).astype(np.uint64) & m
    y = np.asarray(y).astype(np.int64).astype(np.uint64) & m
    z = np.asarray(z).astype(np.int64).astype(np.uint64) & m
    return (x << np.uint64(42)) | (y << np.uint64(21)) | z


def hash_uniform(seed: int, stream: int, x, y, z) -> np.ndarray:
    """Uniform [0, 1) float32 with 24 random bits, a pure function of (seed, stream, x, y, z)."""
    with np.errstate(over="ignore"):
        base = splitmix64(np.uint64(seed) * np.uint64(0x100000001B3) ^ np.uint64(stream))
        h = splitmix64(base ^ _pack(x, y, z))
    return ((h >> np.uint64(40)).astype(np.float64) * (1.0 / 16777216.0)).astype(np.float32)


# --------------------------------------------------------------------------------------
# leaf sets
# --------------------------------------------------------------------------------------
def leaves_touched(inside, lo, hi, chunk: int = 4096) -> np.ndarray:
    """Origins (L, 3) of every 8^3 leaf in the box [lo, hi) holding a voxel with inside(x,y,z)."""
    lo = np.asarray(lo) // 8 * 8
    hi = -(-np.asarray(hi) // 8) * 8
    gx, gy, gz = [np.arange(a, b, 8, dtype=np.int64) for a, b in zip(lo, hi)]
    org = np.stack(np.meshgrid(gx, gy, gz, indexing="ij"), -1).reshape(-1, 3)
    loc = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij"), -1).reshape(-1, 3)
    keep = np.zeros(len(org), dtype=bool)
    for s in range(0, len(org), chunk):
        c = org[s:s + chunk, None, :] + loc[None, :, :]
        keep[s:s + chunk] = inside(c[..., 0], c[..., 1], c[..., 2]).any(axis=1)
    return org[keep].astype(np.int32)


def sphere(cx, cy, cz, r):
    return lambda x, y, z: (x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 <= r * r


def y_cone(cx, cz, y0, height, r0):
    """Vertical (+y) cone: base radius r0 at y0 shrinking linearly to 0 at y0 + height."""

    def f(x, y, z):
        t = (y - y0) / float(height)
        r = r0 * (1.0 - t)
        return (t >= 0) & (t <= 1) & ((x - cx) ** 2 + (z - cz) ** 2 <= r * r)

    return f


def y_cylinder(cx, cz, y0, height, r0):
    return lambda x, y, z: (y >= y0) & (y < y0 + height) & ((x - cx) ** 2 + (z - cz) ** 2 <= r0 * r0)


# --------------------------------------------------------------------------------------
# curl noise
# --------------------------------------------------------------------------------------
def _interp_matrix(n: int, origin: int, spacing: float):
    """1-D smoothstep value-noise weights: rows = voxel positions origin..origin+n-1."""
    pos = (origin + np.arange(n)) / spacing
    i0 = np.floor(pos).astype(np.int64)
    t = pos - i0
    s = t * t * (3.0 - 2.0 * t)
    lat_lo = int(i0.min())
    lat_n = int(i0.max()) - lat_lo + 2
    W = np.zeros((n, lat_n))
    W[np.arange(n), i0 - lat_lo] = 1.0 - s
    W[np.arange(n), i0 - lat_lo + 1] = s
    return W, lat_lo, lat_n


def value_noise_dense(seed: int, stream: int, lo, shape, spacing: float) -> np.ndarray:
    """Smooth value noise in [-1, 1] on the dense voxel box lo + [0, shape), separable evaluation."""
    Ws = [_interp_matrix(n, int(o), spacing) for n, o in zip(shape, lo)]
    (Wx, lx, nx), (Wy, ly, ny), (Wz, lz, nz) = Ws
    gi = np.arange(lx, lx + nx)
    gj = np.arange(ly, ly + ny)
    gk = np.arange(lz, lz + nz)
    I, J, K = np.meshgrid(gi, gj, gk, indexing="ij")
    lat = hash_uniform(seed, stream, I, J, K).astype(np.float64) * 2.0 - 1.0
    a = np.tensordot(Wx, lat, axes=(1, 0))           # (sx, ny, nz)
    a = np.tensordot(a, Wy, axes=(1, 1))             # (sx, nz, sy)
    a = np.tensordot(a, Wz, axes=(1, 1))             # (sx, sy, sz)
    return a.astype(np.float32)


def curl_noise(coords: np.ndarray, seed: int, octaves: int, base_wavelength: float = 32.0,
               amplitude: float = 1.0) -> np.ndarray:
    """Divergence-free-ish velocity (N, 3) float32, max |u| over coords scaled to `amplitude`."""
    lo = coords.min(axis=0).astype(np.int64) - 1
    hi = coords.max(axis=0).astype(np.int64) + 2
    shape = tuple(int(v) for v in (hi - lo))
    psi = [np.zeros(shape, np.float32) for _ in range(3)]
    for o in range(octaves):
        lam = base_wavelength / (2 ** o)
        amp = 0.5 ** o
        for c in range(3):
            psi[c] += amp * value_noise_dense(seed, 16 * o + c, lo, shape, lam)
    ix = coords[:, 0] - lo[0]
    iy = coords[:, 1] - lo[1]
    iz = coords[:, 2] - lo[2]

    def d(f, axis):
        if axis == 0:
            return 0.5 * (f[ix + 1, iy, iz] - f[ix - 1, iy, iz])
        if axis == 1:
            return 0.5 * (f[ix, iy + 1, iz] - f[ix, iy - 1, iz])
        return 0.5 * (f[ix, iy, iz + 1] - f[ix, iy, iz - 1])

    u = np.stack([d(psi[2], 1) - d(psi[1], 2), d(psi[0], 2) - d(psi[2], 0), d(psi[1], 0) - d(psi[0], 1)], -1)
    m = float(np.sqrt((u.astype(np.float64) ** 2).sum(-1)).max())
    if m > 0:
        u = u * (amplitude / m)
    return u.astype(np.float32)


# --------------------------------------------------------------------------------------
# scenes
# --------------------------------------------------------------------------------------
@dataclasses.dataclass
class Scene:
    name: str
    coords: np.ndarray                     # (N, 3) int32, leaf-dense, canonical leaf order
    voxel_size: float
    dt: float
    iterations: int
    params: Dict[str, float]               # CombustionParams fields
    velocity: np.ndarray                   # (N, 3) float32
    fields: Dict[str, np.ndarray]          # state scalars (float32, N)
    sources: Dict[str, np.ndarray] = dataclasses.field(default_factory=dict)
    collision_sdf: Optional[np.ndarray] = None
    notes: str = ""

    @property
    def n(self) -> int:
        return int(self.coords.shape[0])

    @property
    def n_leaves(self) -> int:
        return self.n // 512


def default_params(**kw) -> Dict[str, float]:
    """CombustionParams defaults of the HNanoSolver SOP (SOP_HNanoSolver.cpp:76-167)."""
    p = dict(expansionRate=0.1, temperatureRelease=0.5, buoyancyStrength=1.0, ambientTemp=23.0,
             vorticityScale=1.0, factorScale=0.5, disturbanceEnable=1.0, disturbanceStrength=0.0,
             disturbanceSwirl=0.0, disturbanceThreshold=0.0, disturbanceGain=0.0, disturbanceFrequency=0.0,
             substeps=1, gravity=1.0, maskDensityMin=0.0, maskDensityMax=1.0, downloadVel=1, simReset=0)
    p.update(kw)
    return p


def _coords(origins: np.ndarray) -> np.ndarray:
    return coords_from_leaves(origins[leaf_order(origins)])


def smoke64(seed: int = 1) -> Scene:
    """Plain smoke step on a 64^3 box: leaves touched by the radius-24 ball, Gaussian density,
    solid-body rotation about z (omega 0.5) plus U[-0.1, 0.1] noise, dt 0.1, dx 0.1, K 60."""
    dx = 0.1
    org = leaves_touched(sphere(32, 32, 32, 24), (0, 0, 0), (64, 64, 64))
    c = _coords(org)
    x, y, z = [c[:, i].astype(np.float64) for i in range(3)]
    r2 = (x - 32) ** 2 + (y - 32) ** 2 + (z - 32) ** 2
    dens = np.exp(-r2 / (2.0 * 8.0 ** 2)).astype(np.float32)
    vel = np.stack([-0.5 * (y - 32) * dx, 0.5 * (x - 32) * dx, np.zeros_like(x)], -1)
    noise = np.stack([hash_uniform(seed, 100 + k, c[:, 0], c[:, 1], c[:, 2]) for k in range(3)], -1)
    vel = (vel + (noise.astype(np.float64) * 0.2 - 0.1)).astype(np.float32)
    params = default_params(disturbanceEnable=0.0, gravity=0.0)
    return Scene("smoke64", c, dx, 0.1, 60, params, vel, {"density": dens},
                 notes="leaf-dense N=94208 (184 leaves); gravity 0 and disturbance off so K is literal")


def _fire(name, origins, seed, octaves, dx, dt, K, fuel_pred, notes=""):
    c = _coords(origins)
    vel = curl_noise(c, seed, octaves, 32.0, 1.0)
    n = len(c)
    fuel = np.zeros(n, np.float32)
    m = fuel_pred(c[:, 0], c[:, 1], c[:, 2])
    fuel[m] = hash_uniform(seed, 7, c[m, 0], c[m, 1], c[m, 2])
    fields = {"density": np.zeros(n, np.float32), "temperature": np.zeros(n, np.float32), "fuel": fuel}
    params = default_params(vorticityScale=0.2, factorScale=1.0, temperatureRelease=0.9, ambientTemp=0.0,
                            buoyancyStrength=1.0, gravity=0.0, disturbanceEnable=0.0, substeps=1,
                            expansionRate=0.1)
    return Scene(name, c, dx, dt, K, params, vel, fields, notes=notes)


def fire256_cone(seed: int = 2) -> Scene:
    """Fire step, 256^3 box, +y cone of base radius 80 over 240 voxels (the literal geometry)."""
    shape = y_cone(128, 128, 8, 240, 80.0)
    org = leaves_touched(shape, (0, 0, 0), (256, 256, 256))
    fuel = lambda x, y, z: shape(x, y, z) & (y < 8 + 16)
    return _fire("fire256_cone", org, seed, 3, 0.1, 0.05, 100, fuel,
                 notes="literal cone: leaf-dense N about 1.98M")


def fire256_8m(seed: int = 2) -> Scene:
    """Fire step at the 8M-voxel production size: 256^3 box, +y cylinder of radius 101 over 240
    voxels (a cone cannot hold 8M voxels in 256^3), fuel U[0,1] in its bottom 16 layers.
    Leaf-dense N = 8,355,840 (16,320 leaves)."""
    shape = y_cylinder(128, 128, 8, 240, 101.0)
    org = leaves_touched(shape, (0, 0, 0), (256, 256, 256))
    fuel = lambda x, y, z: shape(x, y, z) & (y < 8 + 16)
    return _fire("fire256_8m", org, seed, 3, 0.1, 0.05, 100, fuel,
                 notes="y-cylinder R=101 H=240 in 256^3")


def fire512(seed: int = 5) -> Scene:
    """Large fire: 512^3 box, cone scaled x2 (base radius 160 over 480 voxels), 5-octave noise, K 150."""
    shape = y_cone(256, 256, 16, 480, 160.0)
    org = leaves_touched(shape, (0, 0, 0), (512, 512, 512))
    fuel = lambda x, y, z: shape(x, y, z) & (y < 16 + 32)
    return _fire("fire512", org, seed, 5, 0.1, 0.05, 150, fuel, notes="cone x2 in 512^3")


def blobs(coords: np.ndarray, seed: int, n_blobs: int, sigma: float, centre, radius: float) -> np.ndarray:
    """Sum of seeded Gaussian blobs with centres uniform in a ball."""
    rng_c = []
    k = 0
    while len(rng_c) < n_blobs:
        u = np.array([hash_uniform(seed, 1000 + k, i, 0, 0) for i in range(3)], np.float64) * 2 - 1
        k += 1
        if (u ** 2).sum() <= 1.0:
            rng_c.append(np.asarray(centre, np.float64) + radius * u)
    x = coords.astype(np.float64)
    out = np.zeros(len(coords))
    for cc in rng_c:
        out += np.exp(-((x - cc) ** 2).sum(-1) / (2 * sigma * sigma))
    return out.astype(np.float32)


def control128(seed: int = 3, target_seed: int = 33) -> Scene:
    """Differentiable density control: 128^3 box, ball of radius 60, 8 blobs (sigma 6)."""
    org = leaves_touched(sphere(64, 64, 64, 60), (0, 0, 0), (128, 128, 128))
    c = _coords(org)
    dens = blobs(c, seed, 8, 6.0, (64, 64, 64), 40.0)
    target = blobs(c, target_seed, 8, 6.0, (64, 64, 64), 40.0)
    vel = curl_noise(c, seed, 3, 32.0, 1.0)
    params = default_params(disturbanceEnable=0.0, gravity=0.0, vorticityScale=0.0)
    sc = Scene("control128", c, 0.1, 0.05, 80, params, vel, {"density": dens},
               notes="16 steps, RBGS 80, checkpoint every 4")
    sc.sources["target_density"] = target
    return sc


def control64(seed: int = 3, target_seed: int = 33) -> Scene:
    """The density-control problem at test size: 64^3 box, ball of radius 28, 8 blobs (sigma 4)."""
    org = leaves_touched(sphere(32, 32, 32, 28), (0, 0, 0), (64, 64, 64))
    c = _coords(org)
    dens = blobs(c, seed, 8, 4.0, (32, 32, 32), 18.0)
    target = blobs(c, target_seed, 8, 4.0, (32, 32, 32), 18.0)
    vel = curl_noise(c, seed, 3, 16.0, 1.0)
    params = default_params(disturbanceEnable=0.0, gravity=0.0, vorticityScale=0.0)
    sc = Scene("control64", c, 0.1, 0.05, 80, params, vel, {"density": dens},
               notes="16 steps, RBGS 80, checkpoint every 4")
    sc.sources["target_density"] = target
    return sc


def precond128(seed: int) -> Scene:
    """Learned-preconditioner scenes: the fire scene at 128^3 (cone of base radius 40 over 120)."""
    shape = y_cone(64, 64, 4, 120, 40.0)
    org = leaves_touched(shape, (0, 0, 0), (128, 128, 128))
    fuel = lambda x, y, z: shape(x, y, z) & (y < 4 + 8)
    return _fire("precond128", org, seed, 3, 0.1, 0.05, 100, fuel)


def precond128_2m(seed: int) -> Scene:
    """Learned-preconditioner training and evaluation scenes at about 2M voxels in 128^3: a +y
    cylinder whose radius (48..62) and height (104..120) vary with the seed, fire fields and
    3-octave curl noise as in the fire scene."""
    u = hash_uniform(seed, 900, np.array([0, 1]), np.array([0, 0]), np.array([0, 0]))
    r0 = 48.0 + 14.0 * float(u[0])
    h = 104 + int(16 * float(u[1]))
    shape = y_cylinder(64, 64, 4, h, r0)
    org = leaves_touched(shape, (0, 0, 0), (128, 128, 128))
    fuel = lambda x, y, z: shape(x, y, z) & (y < 4 + 8)  # noqa: E731
    return _fire("precond128_2m", org, seed, 3, 0.1, 0.05, 100, fuel,
                 notes="cylinder R=%.1f H=%d in 128^3" % (r0, h))


SCENES = {"smoke64": smoke64, "fire256_cone": fire256_cone, "fire256_8m": fire256_8m, "fire512": fire512,
          "control128": control128, "control64": control64}
