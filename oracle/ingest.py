"""TEST INFRASTRUCTURE — brute-force restatement of the SOP's host ingestion and output scatter.

Only tests/ import this module. It restates, voxel by voxel and for small inputs only:
  - the domain of SOP_HNanoSolver.cpp:276-287: the primary velocity grid's active voxels dilated by
    `padding` (openvdb dilateVoxels with NN_FACE_EDGE_VERTEX: each pass adds the 26 neighbours, so
    `padding` passes reach every voxel within L-infinity distance `padding`), united with the
    collision SDF's active voxels;
  - IndexGridBuilder::computeLeafs / build (src/Utils/GridBuilder.hpp:93-175, :231-251): every voxel
    of every domain leaf, leaves in tree order (the oracle's own tree traversal, oracle.leaf_order),
    one block per grid name in insertion order (GridIndexedData::addValueBlock refuses a repeated
    name, GridData.hpp:50-53, so the first grid of a name is the one whose values land), each block's
    leaf copied from the grid's leaf at the same origin (probeConstLeaf + memcpy of the buffer, active
    or not), else memset 0, and for the SDF memset(..., 1, ...): every byte 0x01;
  - the output scatter of SOP_HNanoSolver.cpp:345-388: setValueOn(coord, value) for each voxel whose
    mask byte is set, on a grid whose other voxels hold the background;
  - IndexGridBuilder::writeIndexGrid (GridBuilder.hpp:177-226): a grid whose topology is the domain
    grid's (topologyUnion: its active voxels are the domain's) and whose leaf buffers are memcpy'd from
    the payload, every payload leaf in payload order.
Grids are dicts {"name", "origins" (L, 3), "values" (L, 512[, 3]), "active" (L, 512) bool, "sdf"}.
"""
from __future__ import annotations

import itertools

import numpy as np

from . import leaf_order

LOCAL = np.array([(x, y, z) for x in range(8) for y in range(8) for z in range(8)], np.int64)
SDF_FILL = np.frombuffer(b"\x01\x01\x01\x01", np.float32)[0]  # the float memset(..., 1, ...) leaves


def active_voxels(g):
    out = []
    for l, o in enumerate(g["origins"]):
        for k in range(512):
            if g["active"][l, k]:
                out.append(tuple(int(v) for v in (np.asarray(o, np.int64) + LOCAL[k])))
    return out


def domain_voxels(grids, domain, padding):
    """The domain's active voxels: grids[domain]'s dilated by `padding` (every voxel within
    L-infinity distance `padding`), united with the first SDF grid's."""
    vox = set()
    r = range(-padding, padding + 1)
    for c in active_voxels(grids[domain]):
        for d in itertools.product(r, r, r):
            vox.add((c[0] + d[0], c[1] + d[1], c[2] + d[2]))
    for g in grids:
        if g.get("sdf"):
            vox.update(active_voxels(g))
            break
    return vox


def domain_leaves(grids, domain, padding):
    """Leaf origins of the dilated domain in tree order."""
    vox = domain_voxels(grids, domain, padding)
    leaves = sorted({(x >> 3 << 3, y >> 3 << 3, z >> 3 << 3) for x, y, z in vox})
    org = np.asarray(leaves, np.int32).reshape(-1, 3)
    return org[leaf_order(org)] if len(org) else org


def build(grids, domain, padding):
    """(coords (N, 3), {name: values}) of the payload, blocks in insertion order."""
    org = domain_leaves(grids, domain, padding)
    coords = (org[:, None, :].astype(np.int64) + LOCAL[None]).reshape(-1, 3).astype(np.int32)
    blocks = {}
    for g in grids:
        if g["name"] in blocks:
            continue
        vec3 = np.asarray(g["values"]).ndim == 3
        at = {tuple(int(v) for v in o): l for l, o in enumerate(g["origins"])}
        out = []
        for o in org:
            l = at.get(tuple(int(v) for v in o))
            if l is not None:
                out.append(np.asarray(g["values"][l], np.float32))
            else:
                fill = SDF_FILL if g.get("sdf") else np.float32(0.0)
                out.append(np.full((512, 3) if vec3 else (512,), fill, np.float32))
        blocks[g["name"]] = np.concatenate(out) if out else np.zeros((0, 3) if vec3 else 0, np.float32)
    return coords, blocks


def scatter(coords, values, mask, background):
    """{coord: value} of the set-on voxels and the leaf origins that hold them, in payload order."""
    on = {}
    leaves = []
    for i in range(len(coords)):
        if mask is not None and not mask[i]:
            continue
        c = tuple(int(v) for v in coords[i])
        on[c] = values[i]
        o = (c[0] >> 3 << 3, c[1] >> 3 << 3, c[2] >> 3 << 3)
        if not leaves or leaves[-1] != o:
            if o not in leaves:
                leaves.append(o)
    return on, leaves


def write_index_grid(coords, values, vox):
    """(origins, leaf buffers, active (L, 512) bool) of writeIndexGrid: every payload leaf, its 512
    values copied, a voxel active iff it is a domain voxel (vox: the set of domain_voxels)."""
    L = len(coords) // 512
    origins = [tuple(int(v) for v in coords[512 * l]) for l in range(L)]
    active = np.array([[tuple(int(v) for v in coords[512 * l + o]) in vox for o in range(512)] for l in range(L)])
    buf = np.asarray(values).reshape((L, 512, 3) if np.asarray(values).ndim == 2 else (L, 512))
    return origins, buf.copy(), active
