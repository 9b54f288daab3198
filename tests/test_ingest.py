"""CPU checks of the host ingestion (hns_payload_build), the domain's voxel masks
(hns_payload_domain_mask), the output scatter (hns_scatter_leaves) and the writeIndexGrid export
(hns_write_leaf_grid) against the brute-force restatement in oracle/ingest.py (the SOP's dilated
domain and IndexGridBuilder, SOP_HNanoSolver.cpp:276-301, GridBuilder.hpp:93-226; the scatter of
SOP_HNanoSolver.cpp:345-388), bit for bit, on small random sparse grids."""
import numpy as np
import pytest

import hnanosolver_amd as hns
from hnanosolver_amd import ingest
from oracle import ingest as ref

BOUNDARY_ORIGINS = [(0, 0, 0), (-8, 0, 8), (120, -8, 0), (128, 0, -8), (4088, 8, 8), (-4096, 16, 0), (-136, -136, 128),
                    (8, 8, 8)]


def random_grid(rng, name, origins, density, vec3=False, sdf=False):
    L = len(origins)
    active = rng.random((L, 512)) < density
    values = rng.standard_normal((L, 512, 3) if vec3 else (L, 512)).astype(np.float32)
    return {"name": name, "origins": np.asarray(origins, np.int32), "values": values, "active": active, "sdf": sdf}


def to_leafgrid(g):
    bits = np.zeros((len(g["origins"]), 8), np.uint64)
    for w in range(8):
        chunk = g["active"][:, 64 * w:64 * (w + 1)].astype(np.uint64)
        bits[:, w] = (chunk << np.arange(64, dtype=np.uint64)[None, :]).sum(1, dtype=np.uint64)
    return ingest.LeafGrid(g["name"], g["origins"], g["values"], bits, ingest.ROLE_SDF if g.get("sdf") else 0)


def scene(seed):
    rng = np.random.default_rng(seed)
    vel = random_grid(rng, "vel", BOUNDARY_ORIGINS[:6], 0.02, vec3=True)
    dens = random_grid(rng, "density", [(0, 0, 0), (128, 0, -8), (-4096, 16, 0), (512, 512, 512)], 0.5)
    sdf = random_grid(rng, "collision_sdf", [(-8, 0, 8), (-64, -64, -64), (4096, 8, 8)], 0.05, sdf=True)
    src = random_grid(rng, "src_density", [(120, -8, 0), (8, 8, 8)], 0.3)
    dup = random_grid(rng, "density", [(0, 0, 0)], 1.0)  # a repeated name: the first grid keeps it
    return [dens, vel, sdf, src, dup], 1


@pytest.mark.parametrize("seed,padding", [(0, 0), (1, 1), (2, 2), (3, 9)])
def test_payload_matches_brute_force(seed, padding):
    grids, dom = scene(seed)
    coords, blocks = ref.build(grids, dom, padding)
    p = ingest.build_payload([to_leafgrid(g) for g in grids], dom, padding)
    d = p.data
    print("  padding %d: %d leaves, blocks %s, pinned %s" % (padding, len(coords) // 512, list(d.blocks), p.pinned))
    np.testing.assert_array_equal(d.coords, coords)
    assert list(d.blocks) == list(blocks)
    for k in blocks:
        # byte-for-byte, so the SDF's 0x01 fill and every copied value (NaN-free by construction) match
        assert d[k].tobytes() == blocks[k].tobytes(), k
    p.close()


def test_sdf_fill_is_the_memset_one_pattern():
    grids, dom = scene(5)
    p = ingest.build_payload([to_leafgrid(g) for g in grids], dom, 1)
    sdf = p.data["collision_sdf"].view(np.uint32)
    have = {tuple(o) for o in grids[2]["origins"]}
    L = len(p.data.coords) // 512
    for l in range(L):
        o = tuple(int(v) for v in p.data.coords[512 * l])
        leaf = sdf[512 * l:512 * (l + 1)]
        if o not in have:
            assert (leaf == 0x01010101).all()
    assert np.float32(ref.SDF_FILL) == np.float32(2.3694278e-38)
    p.close()


def test_scatter_matches_brute_force():
    grids, dom = scene(6)
    coords, blocks = ref.build(grids, dom, 2)
    p = ingest.build_payload([to_leafgrid(g) for g in grids], dom, 2)
    rng = np.random.default_rng(3)
    for name in ("density", "vel"):
        for mask in (None, (rng.random(len(coords)) < 0.01).astype(np.uint8)):
            out = ingest.scatter_leaves(p.data, name, mask, background=-7.5)
            on, leaves = ref.scatter(coords, blocks[name], mask, -7.5)
            assert [tuple(o) for o in out.origins] == leaves
            c, v = out.voxels()
            got = {tuple(int(t) for t in cc): vv for cc, vv in zip(c, v)}
            assert got.keys() == on.keys()
            for key, val in on.items():
                np.testing.assert_array_equal(got[key], val)
            # unset voxels of the output leaves hold the background
            full = out.values.reshape(-1, 3) if out.vec3 else out.values.reshape(-1)
            assert np.count_nonzero(np.all(full == -7.5, axis=-1) if out.vec3 else full == -7.5) >= \
                out.values.shape[0] * 512 - len(on)
    p.close()


def test_payload_errors():
    grids, dom = scene(7)
    lg = [to_leafgrid(g) for g in grids]
    with pytest.raises(hns.HNSError) as e:
        ingest.build_payload(lg, dom, -1)
    assert e.value.status == -1
    with pytest.raises(hns.HNSError) as e:
        ingest.build_payload(lg, 2, 1)  # the SDF cannot be the domain grid
    assert e.value.status == -1
    bad = ingest.LeafGrid("x", [(3, 0, 0)], np.zeros((1, 512), np.float32))
    with pytest.raises(hns.HNSError) as e:
        ingest.build_payload([bad], 0, 0)
    assert e.value.status == -4
    rep = ingest.LeafGrid("x", [(0, 0, 0), (0, 0, 0)], np.zeros((2, 512), np.float32))
    with pytest.raises(hns.HNSError) as e:
        ingest.build_payload([rep], 0, 0)
    assert e.value.status == -4
    empty = ingest.LeafGrid("x", [(0, 0, 0)], np.zeros((1, 512), np.float32), np.zeros((1, 8), np.uint64))
    with pytest.raises(hns.HNSError) as e:
        ingest.build_payload([empty], 0, 3)
    assert e.value.status == -2


def bits_to_bool(words):
    w = np.asarray(words, np.uint64).reshape(-1, 8)
    return ((w[:, np.arange(512) >> 6] >> (np.arange(512) & 63).astype(np.uint64)) & np.uint64(1)).astype(bool)


@pytest.mark.parametrize("seed,padding", [(0, 0), (1, 1), (2, 2), (3, 9), (4, 17)])
def test_domain_mask_matches_brute_force(seed, padding):
    """The voxel-level domain (dilation by unit 26-neighbourhood steps, each as three axis shifts of
    64-bit words with carries between leaves) against every voxel within L-infinity distance
    `padding` of an active voxel, plus the SDF's voxels; padding 9 and 17 carry across two and three
    leaves."""
    grids, dom = scene(seed)
    p = ingest.build_payload([to_leafgrid(g) for g in grids], dom, padding)
    got = bits_to_bool(p.domain_mask())
    vox = ref.domain_voxels(grids, dom, padding)
    coords = p.data.coords
    want = np.array([tuple(int(v) for v in c) in vox for c in coords]).reshape(-1, 512)
    print("  padding %d: %d leaves, %d domain voxels" % (padding, len(want), int(want.sum())))
    np.testing.assert_array_equal(got, want)
    assert want.any(axis=1).all()  # every payload leaf holds a domain voxel
    p.close()


def test_write_leaf_grid_matches_and_round_trips():
    grids, dom = scene(8)
    lg = [to_leafgrid(g) for g in grids]
    p = ingest.build_payload(lg, dom, 2)
    d = p.data
    dm = p.domain_mask()
    vox = ref.domain_voxels(grids, dom, 2)
    for name in ("density", "vel", "collision_sdf"):
        out = ingest.write_leaf_grid(d, name, dm)
        org, buf, act = ref.write_index_grid(d.coords, d[name], vox)
        assert [tuple(o) for o in out.origins] == org
        assert out.values.tobytes() == buf.astype(np.float32).tobytes(), name
        np.testing.assert_array_equal(bits_to_bool(out.active), act)
    full = ingest.write_leaf_grid(d, "density")
    assert (full.active == ~np.uint64(0)).all()
    # the exported grids rebuild the same payload: the velocity's active voxels are the domain
    vel, dens = ingest.write_leaf_grid(d, "vel", dm), ingest.write_leaf_grid(d, "density", dm)
    p2 = ingest.build_payload([dens, vel], 1, 0)
    np.testing.assert_array_equal(p2.data.coords, d.coords)
    np.testing.assert_array_equal(p2.domain_mask(), dm)
    for name in ("vel", "density"):
        assert p2.data[name].tobytes() == d[name].tobytes()
    p2.close()
    p.close()

