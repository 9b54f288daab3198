"""CPU pins of the leaf order and index mapping (row a1) and of the samplers (row a2).

The order is pinned by an explicit tree traversal written here from OpenVDB's documented
structure (root table sorted by tile origin, internal nodes visiting children in x-major offset
order, 5-4-3 tree), which is how the reference's IndexGridBuilder orders leaves
(src/Utils/GridBuilder.hpp:164-174 via LeafManager). The three implementations under test (the
oracle, the numpy helper, and libhns' host hns_leaf_order) all use the closed-form voxelsToGrid
key instead (PointsToGrid.cuh:596-651)."""
import json
import os

import numpy as np
import pytest

import oracle
from hnanosolver_amd import topology as topo_np

from _helpers import canonical_coords, ragged_origins

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def tree_traversal_order(origins):
    """Leaf order of a 5-4-3 VDB tree, by building the tree and walking it."""
    tree = {}
    for i, (x, y, z) in enumerate(origins):
        root = (x >> 12, y >> 12, z >> 12)             # 4096^3 root tiles, keyed by origin
        up = ((x & 4095) >> 7, (y & 4095) >> 7, (z & 4095) >> 7)
        lo = ((x & 127) >> 3, (y & 127) >> 3, (z & 127) >> 3)
        tree.setdefault(root, {}).setdefault(up, {})[lo] = i
    order = []
    for root in sorted(tree):                          # std::map<Coord> order: signed lexicographic
        upper = tree[root]
        for up in sorted(upper, key=lambda c: (c[0] << 10) | (c[1] << 5) | c[2]):
            lower = upper[up]
            for lo in sorted(lower, key=lambda c: (c[0] << 8) | (c[1] << 4) | c[2]):
                order.append(lower[lo])
    return np.asarray(order, np.int64)


def _origin_sets():
    rng = np.random.default_rng(0)
    sets = [ragged_origins(0), ragged_origins(1)]
    # random leaves around every node boundary: 0, +-128, +-4096, and far negatives
    for centre in (0, 128, -128, 4096, -4096, 2 ** 20, -(2 ** 20)):
        pts = rng.integers(-24, 24, size=(300, 3)) * 8 + centre
        sets.append(np.unique(pts.astype(np.int32), axis=0))
    sets.append(np.asarray([[0, 0, 0]], np.int32))
    sets.append(np.asarray([[-8, -8, -8], [0, 0, 0], [-8, 0, 0], [0, -8, 0]], np.int32))
    return sets


@pytest.mark.parametrize("k", range(11))
def test_leaf_order_matches_tree_traversal(k):
    org = _origin_sets()[k]
    perm = rng_perm = np.random.default_rng(k).permutation(len(org))
    shuffled = org[rng_perm]
    want = shuffled[tree_traversal_order(shuffled)]
    np.testing.assert_array_equal(shuffled[topo_np.leaf_order(shuffled)], want)
    np.testing.assert_array_equal(shuffled[oracle.leaf_order(shuffled)], want)
    del perm


@pytest.mark.parametrize("k", [0, 2, 5])
def test_libhns_host_leaf_order(k):
    """hns_leaf_order is a host function of the shipped library (no device needed)."""
    import hnanosolver_amd as hns

    org = _origin_sets()[k]
    shuffled = org[np.random.default_rng(7).permutation(len(org))]
    np.testing.assert_array_equal(shuffled[hns.leaf_order(shuffled)], shuffled[tree_traversal_order(shuffled)])


def test_offsets_are_a_bijection_with_misses_zero():
    c = canonical_coords(ragged_origins(2))
    o = oracle.Oracle(c)
    rng = np.random.default_rng(1)
    for i in rng.choice(len(c), 2000, replace=False):
        assert o.offset(*c[i]) == i + 1
    have = set(map(tuple, c.reshape(-1, 512, 3)[:, 0, :] >> 3))
    for _ in range(500):
        p = rng.integers(-40, 40, size=3)
        inside = tuple(p >> 3) in have
        off = o.offset(*p)
        assert (off > 0) == inside
        if inside:
            np.testing.assert_array_equal(c[off - 1], p)


def test_offsets_follow_key_rank_not_input_order():
    """For a payload whose leaves are not in key order, the reference's IndexGrid still numbers
    voxels by key rank (voxelsToGrid sorts), so offset-1 differs from the array position."""
    org = np.asarray([[8, 0, 0], [0, 0, 0]], np.int32)  # key order is (0,0,0) then (8,0,0)
    c = topo_np.coords_from_leaves(org)
    o = oracle.Oracle(c)
    assert o.offset(0, 0, 0) == 1
    assert o.offset(8, 0, 0) == 513
    assert o.offset(9, 1, 2) == 513 + (1 << 6 | 1 << 3 | 2)


def test_not_leaf_dense_rejected():
    with pytest.raises(ValueError):
        oracle.Oracle(np.zeros((100, 3), np.int32))


def _indexgrid_fixture():
    with open(os.path.join(GOLDEN, "indexgrid_known_answers.json")) as fh:
        return json.load(fh)


def _ramp_payload(fx):
    """Leaf-dense payload of Tests/IndexGrid.cpp's ramps: density(i,0,0) = i, velocity(0,i,0) = (i,i,i)."""
    n = fx["ramp_length"]
    org = np.unique(np.asarray([[i >> 3 << 3, 0, 0] for i in range(n)] + [[0, i >> 3 << 3, 0] for i in range(n)],
                               np.int32), axis=0)
    c = canonical_coords(org)
    dens = np.zeros(len(c), np.float32)
    vel = np.zeros((len(c), 3), np.float32)
    for i in range(n):
        dens[np.flatnonzero((c == [i, 0, 0]).all(1))] = i
        vel[np.flatnonzero((c == [0, i, 0]).all(1))] = (i, i, i)
    return c, dens, vel


def test_indexgrid_known_answers():
    """Known answers of the reference's sampler tests (Tests/IndexGrid.cpp:138-142, :213-223,
    :278-281), evaluated by the oracle's samplers on the same ramps."""
    fx = _indexgrid_fixture()
    c, dens, vel = _ramp_payload(fx)
    o = oracle.Oracle(c)
    for case in fx["nearest_density"]:
        assert o.nearest(dens, *case["ijk"]) == case["value"]
    for case in fx["nearest_velocity_x"]:
        assert o.trilinear_vec(vel, *map(float, case["ijk"]))[0] == case["value"]
    for case in fx["trilinear_density"]:
        assert o.trilinear(dens, *case["xyz"]) == pytest.approx(case["value"], rel=0, abs=1e-6)
    for case in fx["trilinear_velocity_x"]:
        assert o.trilinear_vec(vel, *case["xyz"])[0] == pytest.approx(case["value"], rel=0, abs=1e-6)
    d2, v2 = dens * 2.0, vel * 2.0
    for case in fx["doubled"]:
        if case["field"] == "density":
            assert o.trilinear(d2, *case["xyz"]) == pytest.approx(case["value"], abs=1e-6)
        else:
            assert o.trilinear_vec(v2, *case["xyz"])[0] == pytest.approx(case["value"], abs=1e-6)


def test_trilinear_exact_on_affine_fields():
    c = canonical_coords(ragged_origins(3))
    o = oracle.Oracle(c)
    a = np.array([0.25, -0.5, 0.125])
    f = (c @ a + 3.0).astype(np.float32)
    rng = np.random.default_rng(2)
    inner = c[(np.abs(c) < 20).all(1)]
    hits = 0
    for p in inner[rng.choice(len(inner), 300, replace=False)]:
        q = p + rng.random(3) * 0.999
        if all(o.offset(*(np.floor(q).astype(int) + d)) for d in np.ndindex(2, 2, 2)):
            assert o.trilinear(f, *q) == pytest.approx(float(q @ a + 3.0), abs=1e-5)
            hits += 1
    assert hits > 50


NANOVDB_INDEX = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                             "nanovdb_index")


@pytest.mark.skipif(not os.path.exists(NANOVDB_INDEX),
                    reason="oracle/_ref/nanovdb_index is built only where the reference's NanoVDB headers are")
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_index_matches_nanovdb(seed, tmp_path):
    """The leaf order and voxel indices against NanoVDB itself: a ValueOnIndex grid built by the
    reference's vendored NanoVDB (oracle/_ref/nanovdb_index) from the payload's voxels, given in a
    shuffled order, has its leaves in the order hns_leaf_order (the product's C++) and the oracle's
    tree traversal give, and every voxel's index (the offset the reference's kernels read, a miss
    reading 0) is 1 + its position in the leaf-dense payload."""
    import struct
    import subprocess

    import hnanosolver_amd as hns
    from oracle import leaf_order as oracle_order
    origins = ragged_origins(seed)
    coords = canonical_coords(origins)
    shuffled = coords[np.random.default_rng(seed).permutation(len(coords))]
    with open(tmp_path / "in.bin", "wb") as f:
        f.write(struct.pack("<q", len(shuffled)) + shuffled.astype("<i4").tobytes())
    subprocess.run([NANOVDB_INDEX, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], check=True, timeout=120)
    raw = open(tmp_path / "out.bin", "rb").read()
    L = struct.unpack("<q", raw[:8])[0]
    nano_leaves = np.frombuffer(raw[8:8 + 12 * L], "<i4").reshape(L, 3)
    idx = np.frombuffer(raw[8 + 12 * L:], "<u8")
    assert L == len(origins)
    np.testing.assert_array_equal(nano_leaves, origins[hns.leaf_order(origins)])
    np.testing.assert_array_equal(nano_leaves, origins[oracle_order(origins)])
    np.testing.assert_array_equal(nano_leaves, coords.reshape(-1, 512, 3)[:, 0, :])
    # the index of each shuffled voxel is 1 + its position in the payload
    pos = {tuple(c): i for i, c in enumerate(coords.tolist())}
    want = np.array([pos[tuple(c)] + 1 for c in shuffled.tolist()], np.uint64)
    np.testing.assert_array_equal(idx, want)
    print("  %d leaves, %d voxels: NanoVDB order and indices equal" % (L, len(coords)))
