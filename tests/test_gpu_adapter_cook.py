"""GPU run of the SOP-shaped cook through the C++ adapter (tests/cpp/adapter_cook.cpp over
include/hns.hpp, built in-tree by build()): two cooks of SOP_HNanoSolver.cpp:205-396 on the same
leaf grids with the context and topology cached between them. The first cook builds the topology,
the second keeps it; both cooks' scattered outputs are compared with the FP32 oracle stepped from
the same payload (two chained HNS_Advance calls with the sources uploaded each cook)."""
import os
import struct
import subprocess

import numpy as np
import pytest

from hnanosolver_amd import ingest, scenes
from oracle import Oracle

from _helpers import TOL_REL, rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "adapter_cook")
# the cook's CombustionParams (adapter_cook.cpp): the SOP defaults with gravity 0 and no disturbance
PARAMS = scenes.default_params(disturbanceEnable=0.0, gravity=0.0)


def _grids():
    """A sparse smoke: velocity active within radius 24 of (32, 32, 32), density within radius 18
    (so domain leaves without a density leaf read 0), a small density source."""
    g = np.stack(np.meshgrid(*[np.arange(2, 62)] * 3, indexing="ij"), -1).reshape(-1, 3)
    r2 = ((g - 32) ** 2).sum(1)
    vc = g[r2 <= 24 ** 2]
    x, y = vc[:, 0].astype(np.float64), vc[:, 1].astype(np.float64)
    vel = np.stack([-0.5 * (y - 32) * 0.1, 0.5 * (x - 32) * 0.1, 0.05 * np.sin(0.3 * x)], -1).astype(np.float32)
    dc = g[r2 <= 18 ** 2]
    dens = np.exp(-((dc - 32) ** 2).sum(1) / 128.0).astype(np.float32)
    sc = g[((g - (32, 20, 32)) ** 2).sum(1) <= 16]
    return [ingest.LeafGrid.from_voxels("density", dc, dens), ingest.LeafGrid.from_voxels("vel", vc, vel),
            ingest.LeafGrid.from_voxels("src_density", sc, np.full(len(sc), 0.5, np.float32))], 1


def _write_input(path, grids, dom, dx, dt, K, padding):
    with open(path, "wb") as f:
        f.write(struct.pack("<ffiiii", dx, dt, K, padding, len(grids), dom))
        for g in grids:
            nm = g.name.encode()
            f.write(struct.pack("<i", len(nm)) + nm)
            f.write(struct.pack("<iiq", 3 if g.vec3 else 1, g.role, len(g.origins)))
            f.write(g.origins.astype("<i4").tobytes())
            f.write(g.values.astype("<f4").tobytes())
            act = g.active if g.active is not None else np.full((len(g.origins), 8), ~np.uint64(0), np.uint64)
            f.write(act.astype("<u8").tobytes())


def _read_output(path, grids):
    out = []
    with open(path, "rb") as f:
        for _ in range(2):
            builds, reuses, pinned = struct.unpack("<qqi", f.read(20))
            fields = {}
            for g in grids:
                if g.role != ingest.ROLE_VALUE:
                    continue
                L = struct.unpack("<q", f.read(8))[0]
                comps = 3 if g.vec3 else 1
                org = np.frombuffer(f.read(12 * L), "<i4").reshape(L, 3)
                val = np.frombuffer(f.read(4 * 512 * comps * L), "<f4").reshape((L * 512, 3) if comps == 3 else (-1,))
                f.read(64 * L)
                fields[g.name] = (org, val)
            out.append((builds, reuses, pinned, fields))
    return out


@pytest.mark.timeout(300, method="thread")
def test_adapter_cook_two_cooks_match_oracle(hip, tmp_path):
    assert os.path.exists(EXE), "tests/cpp/_build/adapter_cook is not built (run __graft_entry__.build())"
    grids, dom = _grids()
    dx, dt, K, padding = 0.1, 0.1, 40, 2
    print("CASE adapter cook x2: %s, padding %d, K=%d" % ([(g.name, len(g.origins)) for g in grids], padding, K),
          flush=True)
    _write_input(tmp_path / "in.bin", grids, dom, dx, dt, K, padding)
    r = subprocess.run([EXE, "cook", str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True,
                       text=True, timeout=240)
    print(r.stdout.strip(), r.stderr.strip()[-400:])
    assert r.returncode == 0 and "cook: ok" in r.stdout
    cooks = _read_output(tmp_path / "out.bin", grids)
    assert [(c[0], c[1]) for c in cooks] == [(1, 0), (1, 1)]  # the second cook kept the topology
    assert all(c[2] == 1 for c in cooks)  # the payload was in pinned host memory

    # the oracle from the same payload: two chained calls, the density source uploaded each cook
    p = ingest.build_payload(grids, dom, padding)
    d = p.data
    coords = np.array(d.coords)
    o = Oracle(coords)
    vel = np.array(d["vel"])
    fields = {"density": np.array(d["density"])}
    sources = {"src_density": np.array(d["src_density"])}
    p.close()
    leaves = coords.reshape(-1, 512, 3)[:, 0, :]
    for c, (builds, reuses, pinned, out) in enumerate(cooks):
        vel, fields = o.advance(dx, K, dt, PARAMS, vel, fields, sources)
        for name, ref in (("vel", vel), ("density", fields["density"])):
            org, val = out[name]
            np.testing.assert_array_equal(org, leaves)  # every payload leaf, in payload order (mask = 1)
            e = rel_err(val, ref)
            print("  cook %d %-8s rel %.3e (N=%d)" % (c + 1, name, e, len(coords)), flush=True)
            assert e <= TOL_REL


@pytest.mark.timeout(300, method="thread")
def test_topology_kept_per_cook_and_overhead_at_8m(hip):
    """Per-cook host work at the 8M-voxel fire scene: the payload built from leaf grids, the
    topology kept on an unchanged payload and rebuilt on a changed one, and the context upload from
    the pinned payload against the same arrays in pageable memory (times printed)."""
    import time

    import hnanosolver_amd as hns
    torch = hip
    sc = scenes.fire256_8m()
    L = sc.n // 512
    org = sc.coords.reshape(L, 512, 3)[:, 0, :]
    vel = ingest.LeafGrid("vel", org, sc.velocity.reshape(L, 512, 3))
    grids = [ingest.LeafGrid(k, org, v.reshape(L, 512)) for k, v in sc.fields.items()] + [vel]
    print("CASE per-cook overhead fire256_8m N=%d" % sc.n, flush=True)
    t0 = time.perf_counter()
    p = ingest.build_payload(grids, len(grids) - 1, 0)
    t1 = time.perf_counter()
    d = p.data
    np.testing.assert_array_equal(d.coords, sc.coords)
    topo = hns.create_index_grid(d, sc.voxel_size)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    topo = hns.create_index_grid(d, sc.voxel_size, topo)
    t3 = time.perf_counter()
    assert topo.builds() == (1, 1)
    ctx = hns.ensure_context(None, d, topo, sc.voxel_size)
    t4 = time.perf_counter()
    hns.destroy_context(ctx)
    pageable = hns.GridIndexedData(np.array(d.coords))
    for k, a in d.blocks.items():
        pageable.add_value_block(k, np.array(a))
    t5 = time.perf_counter()
    ctx = hns.ensure_context(None, pageable, topo, sc.voxel_size)
    t6 = time.perf_counter()
    hns.destroy_context(ctx)
    print("  payload build %.1f ms (pinned %s); topology build %.1f ms, kept %.3f ms; context upload from pinned "
          "%.1f ms, from pageable %.1f ms" % (1e3 * (t1 - t0), p.pinned, 1e3 * (t2 - t1), 1e3 * (t3 - t2),
                                               1e3 * (t4 - t3), 1e3 * (t6 - t5)), flush=True)
    assert p.pinned
    # a changed payload (one leaf fewer) is rebuilt
    p2 = ingest.build_payload([ingest.LeafGrid("vel", org[1:], sc.velocity.reshape(L, 512, 3)[1:])], 0, 0)
    topo2 = hns.create_index_grid(p2.data, sc.voxel_size, topo)
    assert topo2.builds() == (1, 0) and topo2.info()[1] == L - 1
    p2.close()
    p.close()
