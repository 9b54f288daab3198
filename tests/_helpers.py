"""Shared helpers of the test suite: parity metrics and edge-case topologies."""
import numpy as np

import hnanosolver_amd as hns
from hnanosolver_amd import scenes
from hnanosolver_amd.topology import coords_from_leaves, leaf_order
from oracle import Oracle

# FP32 parity bar of the step: max |gpu - oracle| <= TOL_REL * max |oracle| (normwise relative)
TOL_REL = 1e-5


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(float(np.abs(b).max()) if b.size else 0.0, 1e-30)
    return float(np.abs(a - b).max()) / scale if a.size else 0.0


def max_abs(a, b):
    return float(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)).max()) if np.size(a) else 0.0


def canonical_coords(origins):
    origins = np.asarray(origins, np.int32).reshape(-1, 3)
    return coords_from_leaves(origins[leaf_order(origins)])


def ragged_origins(seed=0):
    """Irregular leaf set: holes, negative coordinates, and clusters across the x = +-128 and
    +-4096 node boundaries (the upper/root levels of the leaf order)."""
    rng = np.random.default_rng(seed)
    out = []
    g = np.arange(-32, 32, 8)
    for x in g:
        for y in g:
            for z in g:
                if rng.random() < 0.55:
                    out.append((x, y, z))
    for base in ((120, 0, 0), (4088, -8, 0), (-4104, 8, 16), (-136, -136, 128)):
        for dx in (0, 8, 16):
            for dy in (0, 8):
                for dz in (0, 8):
                    if rng.random() < 0.8:
                        out.append((base[0] + dx, base[1] + dy, base[2] + dz))
    return np.unique(np.asarray(out, np.int32), axis=0)


def ragged_scene(seed=0, cfl=1.5):
    """Leaf-dense payload over ragged_origins with smooth fields and a velocity of max CFL `cfl`."""
    c = canonical_coords(ragged_origins(seed))
    x = c.astype(np.float64)
    n = len(c)
    u = np.stack([np.sin(0.11 * x[:, 1] + 0.3) + 0.3 * np.cos(0.07 * x[:, 2]),
                  np.cos(0.09 * x[:, 0]) + 0.2 * np.sin(0.13 * x[:, 2]),
                  0.5 * np.sin(0.05 * x[:, 0] + 0.08 * x[:, 1])], -1)
    noise = np.stack([scenes.hash_uniform(seed, 40 + k, c[:, 0], c[:, 1], c[:, 2]) for k in range(3)], -1)
    u = u + 0.3 * (noise - 0.5)
    dx, dt = 0.1, 0.05
    u *= cfl * dx / dt / np.abs(u).max()
    dens = (0.5 + 0.5 * np.sin(0.2 * x[:, 0]) * np.cos(0.15 * x[:, 1])).astype(np.float32)
    temp = scenes.hash_uniform(seed, 50, c[:, 0], c[:, 1], c[:, 2]) * 2.0
    fuel = scenes.hash_uniform(seed, 51, c[:, 0], c[:, 1], c[:, 2])
    return c, u.astype(np.float32), {"density": dens, "temperature": temp.astype(np.float32),
                                     "fuel": fuel.astype(np.float32)}, dx, dt


def plane_sdf(coords, normal=(0.6, 0.8, 0.0), offset=0.0, dx=0.1):
    """Signed distance (world units) to a plane, so a band of voxels has 0 <= sdf < 0.1."""
    n = np.asarray(normal, np.float64)
    n /= np.linalg.norm(n)
    return ((coords.astype(np.float64) @ n) * dx - offset).astype(np.float32)


# ---- whole-step parity through the C ABI (GPU tests) -----------------------------------------

def _payload(coords, velocity, fields, sources=None, sdf=None, src_vel=None):
    d = hns.GridIndexedData(coords)
    for k, v in fields.items():
        d.add_value_block(k, v.copy())
    d.add_value_block("velocity", velocity.copy())
    if sdf is not None:
        d.add_value_block("collision_sdf", sdf)
    for k, v in (sources or {}).items():
        d.add_value_block(k, v)
    if src_vel is not None:
        d.add_value_block("src_vel", src_vel)
    return d


def _out_payload(coords, names):
    d = hns.GridIndexedData(coords)
    for k in names:
        d.add_value_block(k, np.zeros(len(coords), np.float32))
    d.add_value_block("velocity", np.zeros((len(coords), 3), np.float32))
    return d


def _run_gpu(coords, dx, velocity, fields, params, K, dt, calls, sources=None, sdf=None, src_vel=None):
    data = _payload(coords, velocity, fields, sources, sdf, src_vel)
    topo = hns.create_index_grid(data, dx)
    ctx = hns.ensure_context(None, data, topo, dx, has_collision=sdf is not None)
    hns.upload_sources(ctx, data)
    p = hns.make_params(**params)
    for _ in range(calls):
        hns.advance(ctx, K, dt, p)
    out = _out_payload(coords, list(dict.fromkeys(list(fields) + ["fuel", "waste", "temperature", "flame"])))
    hns.download_for_output(ctx, out)
    res = {k: out[k] for k in out.blocks if k != "velocity" and ctx.has_field(k)}
    hns.destroy_context(ctx)
    return out["velocity"], res


def _run_oracle(coords, dx, velocity, fields, params, K, dt, calls, sources=None, sdf=None, src_vel=None,
                dtype=np.float32):
    o = Oracle(coords, dtype)
    vel, f = velocity, dict(fields)
    for _ in range(calls):
        vel, f = o.advance(dx, K, dt, params, vel, f, sources, src_vel, sdf)
    return vel, f


def _per_step_parity(tag, coords, dx, velocity, fields, params, K, dt, calls, sources=None, sdf=None, src_vel=None):
    """Each GPU call (HNS_Advance) against the FP32 oracle advanced from the same state (the GPU's
    state after the previous call); the FP64 oracle from that state is reported beside it."""
    data = _payload(coords, velocity, fields, sources, sdf, src_vel)
    topo = hns.create_index_grid(data, dx)
    ctx = hns.ensure_context(None, data, topo, dx, has_collision=sdf is not None)
    hns.upload_sources(ctx, data)
    p = hns.make_params(**params)
    o32, o64 = Oracle(coords), Oracle(coords, np.float64)
    vel, f = velocity, dict(fields)
    worst = 0.0
    for c in range(calls):
        hns.advance(ctx, K, dt, p)
        out = _out_payload(coords, list(dict.fromkeys(list(fields) + ["fuel", "waste", "temperature", "flame"])))
        hns.download_for_output(ctx, out)
        gv = out["velocity"].copy()
        gf = {k: out[k].copy() for k in out.blocks if k != "velocity" and ctx.has_field(k)}
        ov, of = o32.advance(dx, K, dt, params, vel, f, sources, src_vel, sdf)
        ref64 = o64.advance(dx, K, dt, params, vel, f, sources, src_vel, sdf)
        assert set(gf) == set(of)
        worst = max(worst, _compare("%s#%d" % (tag, c), gv, gf, ov, of, ref64))
        vel, f = gv, gf
    hns.destroy_context(ctx)
    return worst


def _compare(tag, gv, gf, ov, of, o64=None):
    worst = 0.0
    for k in ["velocity"] + sorted(of):
        g = gv if k == "velocity" else gf[k]
        r = ov if k == "velocity" else of[k]
        e = rel_err(g, r)
        nd = int(np.count_nonzero(np.asarray(g) != np.asarray(r)))
        line = "  %-8s %-12s rel %.3e  max-abs %.3e  differing %d/%d" % (tag, k, e, max_abs(g, r), nd, np.size(r))
        if o64 is not None:
            r64 = o64[0] if k == "velocity" else o64[1][k]
            line += "  | vs FP64: gpu %.3e  oracle-f32 %.3e" % (max_abs(g, r64), max_abs(r, r64))
        print(line, flush=True)
        worst = max(worst, e)
    return worst
