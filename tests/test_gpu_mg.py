"""GPU checks of the multigrid-preconditioned solve (HNS_SOLVER_MGCG): on smoke and control scenes
and on ragged topologies with absent neighbour leaves it reaches the requested relative residual,
checked by the FP64 CPU oracle's residual of the returned pressure, in far fewer iterations than
plain CG; the solve is deterministic (bitwise equal pressure on a repeat)."""
import numpy as np
import pytest

import hnanosolver_amd as hns
from hnanosolver_amd import kernels as K
from hnanosolver_amd import precond as P
from hnanosolver_amd import scenes
import oracle
from oracle import Oracle

from _helpers import canonical_coords

pytestmark = pytest.mark.gpu


def _case(name):
    print("CASE", name, flush=True)


def _topo(coords, dx):
    return hns.create_index_grid(hns.GridIndexedData(coords), dx)


def _dense(torch, topo, cs):
    n = cs.numel()
    out = torch.empty_like(cs)
    K.from_colour_split(topo, cs, out)
    return out.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("name", ["smoke64", "control64"])
def test_mgcg_reaches_tolerance_in_few_iterations(hip, name):
    torch = hip
    sc = getattr(scenes, name)()
    _case("mgcg vs cg, %s N=%d" % (name, sc.n))
    topo = _topo(sc.coords, sc.voxel_size)
    div = P.scene_divergence(sc, topo)
    p_cg, p_mg, p_mg2 = torch.empty_like(div), torch.empty_like(div), torch.empty_like(div)
    st_cg = K.pressure_solve(topo, div, p_cg, sc.voxel_size, solver="cg", tol=1e-5)
    st = K.pressure_solve(topo, div, p_mg, sc.voxel_size, solver="mgcg", tol=1e-5, max_iters=500)
    st2 = K.pressure_solve(topo, div, p_mg2, sc.voxel_size, solver="mgcg", tol=1e-5, max_iters=500)
    d0 = _dense(torch, topo, div)
    r2, d2 = Oracle(sc.coords, np.float64).residual(_dense(torch, topo, p_mg), d0, sc.voxel_size)
    print("  cg %d iterations, mgcg %d iterations (fallback %d), rel %.2e (oracle-checked %.2e)" % (
        st_cg.iterations, st.iterations, st.fallback_iterations, st.rel_residual, np.sqrt(r2 / d2)))
    assert np.sqrt(r2 / d2) <= 1.01e-5
    assert st.fallback_iterations == 0
    assert st.iterations * 3 <= st_cg.iterations
    assert st2.iterations == st.iterations and bool(torch.equal(p_mg, p_mg2))


def test_mgcg_on_ragged_topology(hip):
    """Holes and thin parts: absent neighbour leaves at every level of the hierarchy."""
    torch = hip
    rng = np.random.default_rng(21)
    org = [(8 * x, 8 * y, 8 * z) for x in range(10) for y in range(8) for z in range(9) if rng.random() < 0.7]
    coords = canonical_coords(np.array(org))
    _case("mgcg ragged, %d leaves" % (len(coords) // 512))
    topo = _topo(coords, 0.1)
    b = rng.standard_normal(len(coords)).astype(np.float32)
    div_dense = torch.from_numpy(b).cuda()
    div = torch.empty_like(div_dense)
    K.to_colour_split(topo, div_dense, div)
    p = torch.empty_like(div)
    st_cg = K.pressure_solve(topo, div, p, 0.1, solver="cg", tol=1e-5, max_iters=5000)
    st = K.pressure_solve(topo, div, p, 0.1, solver="mgcg", tol=1e-5, max_iters=500)
    r2, d2 = Oracle(coords, np.float64).residual(_dense(torch, topo, p), b.astype(np.float64), 0.1)
    print("  cg %d iterations, mgcg %d iterations, oracle-checked rel %.2e" % (st_cg.iterations, st.iterations,
                                                                              np.sqrt(r2 / d2)))
    assert np.sqrt(r2 / d2) <= 1.01e-5
    assert st.iterations < st_cg.iterations


def test_vcycle_graph_replays_match_eager_launches(hip):
    """The V-cycle captured in a HIP graph gives bitwise the eagerly launched V-cycle, on every replay,
    with eager work and synchronisations between the replays, and after a change of input."""
    torch = hip
    sc = scenes.smoke64()
    _case("v-cycle graph vs eager, smoke64")
    topo = _topo(sc.coords, sc.voxel_size)
    rng = np.random.default_rng(5)
    for trial in range(2):
        r = torch.from_numpy(rng.standard_normal(sc.n).astype(np.float32)).cuda()
        z_eager = torch.empty_like(r)
        K.mg_apply(topo, r, z_eager, graph=False)
        z = torch.empty_like(r)  # the same buffers on every replay, so the captured graph is reused
        for rep in range(3):
            z.fill_(float("nan"))
            K.mg_apply(topo, r, z, graph=True)
            n = int((z != z_eager).sum())
            print("  trial %d replay %d: differing elements %d of %d" % (trial, rep, n, sc.n))
            assert n == 0
            K.mg_apply(topo, r, torch.empty_like(r), graph=False)  # eager work between replays


@pytest.mark.parametrize("shape,keep", [((10, 8, 9), 0.7), ((16, 16, 12), 0.7), ((24, 24, 24), 0.72),
                                        ((32, 32, 16), 1.0), ((32, 32, 17), 0.95)])
def test_vcycle_variants_match(hip, shape, keep):
    """Equivalent V-cycle kernel sequences give bitwise the same z: the coarsest level solved inside
    one workgroup (LDS-resident, up to 8192 leaves per colour: 1, 4 and 8 slots per lane here, and
    exactly 16384 leaves) or by per-sweep launches (above 16384 leaves both take this path), level 0
    as two fused 12^3-tile passes or as separate sweeps, restriction and prolongation, eager or
    replayed from a graph."""
    torch = hip
    rng = np.random.default_rng(sum(shape))
    nx, ny, nz = shape
    org = [(8 * x, 8 * y, 8 * z) for x in range(nx) for y in range(ny) for z in range(nz) if rng.random() < keep]
    coords = canonical_coords(np.array(org))
    _case("v-cycle variants, %d leaves" % (len(coords) // 512))
    topo = _topo(coords, 0.1)
    r = torch.from_numpy(rng.standard_normal(len(coords)).astype(np.float32)).cuda()
    ref = torch.empty_like(r)
    K.mg_apply(topo, r, ref, graph=False, coarse_launches=True, unfused=True)
    assert bool(torch.isfinite(ref).all())
    for graph, coarse, unfused in [(False, False, False), (False, True, False), (False, False, True), (True, False, False)]:
        z = torch.empty_like(r)
        K.mg_apply(topo, r, z, graph=graph, coarse_launches=coarse, unfused=unfused)
        n = int((z != ref).sum())
        print("  graph %d coarse-launches %d unfused %d: differing elements %d of %d" % (
            graph, coarse, unfused, n, len(coords)))
        assert n == 0


@pytest.mark.parametrize("case", [((10, 8, 9), 0.7), ((24, 24, 24), 0.72), ((32, 32, 17), 0.95), "smoke64"])
def test_vcycle_matches_the_oracle_bitwise(hip, case):
    """z = V(r) of the library (default fused kernels, replayed from a graph, and the separate-launch
    variant) is bitwise the CPU oracle's V-cycle (oracle/hns_oracle_mg.cpp, the same FP32 operation
    order) on ragged topologies, one above 8192 leaves of a colour (the coarsest level by
    launches), and the smoke scene's topology."""
    torch = hip
    if case == "smoke64":
        sc = scenes.smoke64()
        coords, seed = sc.coords, 7
    else:
        shape, keep = case
        rng = np.random.default_rng(sum(shape) + 1)
        nx, ny, nz = shape
        org = [(8 * x, 8 * y, 8 * z) for x in range(nx) for y in range(ny) for z in range(nz) if rng.random() < keep]
        coords, seed = canonical_coords(np.array(org)), sum(shape)
    _case("v-cycle vs oracle, %d leaves" % (len(coords) // 512))
    topo = _topo(coords, 0.1)
    r = np.random.default_rng(seed).standard_normal(len(coords)).astype(np.float32)
    z_ref = oracle.Multigrid(coords).vcycle(r)
    r_cs = torch.empty(len(coords), device="cuda")
    K.to_colour_split(topo, torch.from_numpy(r).cuda(), r_cs)
    for graph, coarse, unfused in [(True, False, False), (False, True, True)]:
        z_cs = torch.empty_like(r_cs)
        K.mg_apply(topo, r_cs, z_cs, graph=graph, coarse_launches=coarse, unfused=unfused)
        z = torch.empty_like(z_cs)
        K.from_colour_split(topo, z_cs, z)
        z = z.cpu().numpy()
        n = int((z.view(np.uint32) != z_ref.view(np.uint32)).sum())
        print("  graph %d coarse-launches %d unfused %d: differing bit patterns %d of %d, max |diff| %.2e" % (
            graph, coarse, unfused, n, len(coords), float(np.abs(z - z_ref).max())))
        assert n == 0


def test_mgcg_step_graph_batches_match_eager_iterations(hip):
    """The fire step with the multigrid solve: iterations replayed four at a time from a captured graph
    (timing off) and launched one by one with the stop flag read after each (timing on, so events can
    bracket them) give bitwise the same state, over steps that reach the graph batches."""
    sc = scenes.control64()
    _case("mgcg step, graph batches vs eager, control64 N=%d" % sc.n)

    def run(timing):
        d = hns.GridIndexedData(sc.coords)
        for k, v in sc.fields.items():
            d.add_value_block(k, v.copy())
        d.add_value_block("velocity", sc.velocity.copy())
        topo = hns.create_index_grid(d, sc.voxel_size)
        ctx = hns.ensure_context(None, d, topo, sc.voxel_size)
        ctx.set_timing(timing)
        p = hns.make_params(**sc.params)
        its = [hns.advance(ctx, 0, sc.dt, p, solver="mgcg", tol=1e-5).iterations for _ in range(2)]
        out = hns.GridIndexedData(sc.coords)
        names = [k for k in sc.fields if ctx.has_field(k)]
        for k in names:
            out.add_value_block(k, np.zeros(sc.n, np.float32))
        out.add_value_block("velocity", np.zeros((sc.n, 3), np.float32))
        hns.download_for_output(ctx, out)
        hns.destroy_context(ctx)
        return its, {k: out[k].copy() for k in names + ["velocity"]}

    its_g, a = run(False)
    its_e, b = run(True)
    print("  iterations per step: graph %s, eager %s" % (its_g, its_e))
    assert its_g == its_e and min(its_g) >= 5  # iterations 4-7 come from the graph
    for k in a:
        n = int(np.sum(a[k] != b[k]))
        print("  %-12s differing elements %d of %d" % (k, n, a[k].size))
        assert n == 0


def test_mgcg_step_capped_solve_projects_once(hip):
    """A fire step whose solve ends at its cap (no stop raised): the projection that each read of the
    stop flag launches gated on the flag returns at once, and the step launches it after the solve
    (with the deferred x update flushed), so graph runs (timing off) and eager iterations (timing on,
    the projection never handed to the solve) leave bitwise the same state."""
    sc = scenes.control64()
    _case("mgcg step capped at 6 iterations, graph vs eager, control64 N=%d" % sc.n)

    def run(timing):
        d = hns.GridIndexedData(sc.coords)
        for k, v in sc.fields.items():
            d.add_value_block(k, v.copy())
        d.add_value_block("velocity", sc.velocity.copy())
        topo = hns.create_index_grid(d, sc.voxel_size)
        ctx = hns.ensure_context(None, d, topo, sc.voxel_size)
        ctx.set_timing(timing)
        p = hns.make_params(**sc.params)
        st = [hns.advance(ctx, 0, sc.dt, p, solver="mgcg", tol=1e-12, max_iters=6) for _ in range(2)]
        out = hns.GridIndexedData(sc.coords)
        names = [k for k in sc.fields if ctx.has_field(k)]
        for k in names:
            out.add_value_block(k, np.zeros(sc.n, np.float32))
        out.add_value_block("velocity", np.zeros((sc.n, 3), np.float32))
        hns.download_for_output(ctx, out)
        hns.destroy_context(ctx)
        return [x.iterations for x in st], [x.rel_residual for x in st], {k: out[k].copy() for k in names + ["velocity"]}

    its_g, rel_g, a = run(False)
    its_e, rel_e, b = run(True)
    print("  iterations per step: graph %s, eager %s; residual %s" % (its_g, its_e, ["%.2e" % r for r in rel_g]))
    assert its_g == its_e == [6, 6]
    assert all(r > 1e-12 for r in rel_g)  # the cap, not the tolerance, ended both solves
    for k in a:
        n = int(np.sum(a[k] != b[k]))
        print("  %-12s differing elements %d of %d" % (k, n, a[k].size))
        assert n == 0


def test_mgcg_capped_solves_apply_every_update(hip):
    """Solves ended by max_iters (no stop; the multigrid solve has no fallback): the x update of an
    iteration is applied by the next iteration's coarse-level launch, the last one by the flush after
    the solve. Each iteration cuts the residual several times over, so the oracle's residual of the
    returned pressure falls by a clear factor from cap to cap (a missing last update would leave it
    where the previous cap did), and the solver's own residual of the same pressure agrees."""
    torch = hip
    sc = scenes.control64()
    _case("mgcg capped at 1..7 iterations, control64 N=%d" % sc.n)
    topo = _topo(sc.coords, sc.voxel_size)
    div = P.scene_divergence(sc, topo)
    d0 = _dense(torch, topo, div)
    o = Oracle(sc.coords, np.float64)
    rels = []
    for cap in range(1, 8):
        p = torch.empty_like(div)
        st = K.pressure_solve(topo, div, p, sc.voxel_size, solver="mgcg", tol=1e-9, max_iters=cap,
                              allow_unconverged=True)
        r2, d2 = o.residual(_dense(torch, topo, p), d0, sc.voxel_size)
        rel = float(np.sqrt(r2 / d2))
        print("  cap %d: %d iterations, rel %.4e (oracle %.4e), fallback %d" % (
            cap, st.iterations, st.rel_residual, rel, st.fallback_iterations))
        assert st.iterations == cap and st.fallback_iterations == 0
        assert abs(st.rel_residual - rel) <= 1e-3 * rel + 1e-9
        rels.append(rel)
    for a, b in zip(rels, rels[1:]):
        assert b < 0.7 * a


def test_mgcg_long_solve_through_replacements(hip):
    """A tolerance below what FP32 iterates reach: the solve runs to its cap through rejected stops (the
    recurrence residual passes, the replaced b - A x does not) and a periodic residual replacement at
    iteration 64, each preceded by completing the deferred x update. The pressure it returns sits at
    the FP32 floor, and the solver's residual of it agrees with the oracle's."""
    torch = hip
    sc = scenes.smoke64()
    _case("mgcg to 1e-9 capped at 72 iterations, smoke64 N=%d" % sc.n)
    topo = _topo(sc.coords, sc.voxel_size)
    div = P.scene_divergence(sc, topo)
    p = torch.empty_like(div)
    st = K.pressure_solve(topo, div, p, sc.voxel_size, solver="mgcg", tol=1e-9, max_iters=72,
                          allow_unconverged=True)
    r2, d2 = Oracle(sc.coords, np.float64).residual(_dense(torch, topo, p), _dense(torch, topo, div),
                                                     sc.voxel_size)
    rel = float(np.sqrt(r2 / d2))
    print("  %d iterations, rel %.4e (oracle %.4e), fallback %d" % (
        st.iterations, st.rel_residual, rel, st.fallback_iterations))
    assert st.iterations == 72 and st.fallback_iterations == 0
    assert rel < 1e-5
    assert abs(st.rel_residual - rel) <= 1e-2 * rel + 1e-9


def test_mgcg_graph_remade_for_new_cap_and_tolerance(hip):
    """Solves on one topology replay the batch graphs captured by the first (every run of iterations
    between two stop-flag reads is one graph per run length, from iteration 0 on); a later solve with
    a larger iteration cap than any before (its per-iteration arrays are reallocated) and another
    tolerance has them made again, and then matches the same solve on a fresh topology bit for bit."""
    torch = hip
    sc = scenes.control64()
    _case("mgcg graph cache across caps and tolerances, control64 N=%d" % sc.n)
    topo = _topo(sc.coords, sc.voxel_size)
    div = P.scene_divergence(sc, topo)
    p1, p2, p3 = torch.empty_like(div), torch.empty_like(div), torch.empty_like(div)
    st1 = K.pressure_solve(topo, div, p1, sc.voxel_size, solver="mgcg", tol=1e-3, max_iters=40)
    st2 = K.pressure_solve(topo, div, p2, sc.voxel_size, solver="mgcg", tol=1e-5, max_iters=4000)
    fresh = _topo(sc.coords, sc.voxel_size)
    st3 = K.pressure_solve(fresh, div, p3, sc.voxel_size, solver="mgcg", tol=1e-5, max_iters=4000)
    print("  tol 1e-3: %d iterations; tol 1e-5 after it: %d, on a fresh topology: %d" % (
        st1.iterations, st2.iterations, st3.iterations))
    assert st1.iterations >= 1  # the first solve ran from a batch graph
    assert st2.iterations == st3.iterations and bool(torch.equal(p2, p3))
    r2, d2 = Oracle(sc.coords, np.float64).residual(_dense(torch, topo, p2), _dense(torch, topo, div), sc.voxel_size)
    assert np.sqrt(r2 / d2) <= 1.01e-5
