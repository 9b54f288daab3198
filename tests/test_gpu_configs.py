"""GPU runs of the fire configurations at their stated workloads, checked against the oracle.

Config 2, the fire step at production size (fire256_cone: the literal cone in 256^3, N = 1,976,832;
fire256_8m: the 8M-voxel plume, N = 8,355,840): whole steps with the reference's fixed 100 RBGS
iterations against the FP32 oracle (the HNS_Advance sequence, HNanoSolver.cu:223-456), and the
to-tolerance solves (multigrid-preconditioned CG, CG) re-checked by the FP64 oracle residual, with
bitwise determinism. Config 5, the large fire (fire512: cone x2 in 512^3, N = 14,300,160): a step with
the reference's 150 RBGS iterations against the FP32 oracle, the multigrid and CG solves to 1e-5
re-checked in FP64, and a 4-step checkpointed adjoint with its device memory reported.

Each check prints the sizes, errors and times it measured; the parity bar is TOL_REL (1e-5
normwise relative, per field and step)."""
import time

import numpy as np
import pytest

import hnanosolver_amd as hns
from hnanosolver_amd import diff
from hnanosolver_amd import kernels as K
from hnanosolver_amd import precond as P
from hnanosolver_amd import scenes
from oracle import Oracle

from _helpers import TOL_REL, _compare, _out_payload, _payload, _per_step_parity

pytestmark = pytest.mark.gpu


def _case(name):
    print("CASE", name, flush=True)


_SCENES = {}


def _scene(name):
    """Scenes are pure functions of their seed: built once per test module."""
    if name not in _SCENES:
        t0 = time.perf_counter()
        _SCENES[name] = scenes.SCENES[name]()
        print("  scene %s built in %.1f s" % (name, time.perf_counter() - t0), flush=True)
    return _SCENES[name]


def _dense(torch, topo, cs):
    out = torch.empty_like(cs)
    K.from_colour_split(topo, cs, out)
    return out.cpu().numpy().astype(np.float64)


def _one_step_parity(tag, sc, K_):
    """One HNS_Advance call with K_ RBGS iterations against the FP32 oracle (FP64 oracle printed)."""
    t0 = time.perf_counter()
    worst = _per_step_parity(tag, sc.coords, sc.voxel_size, sc.velocity, sc.fields, sc.params, K_, sc.dt, 1)
    print("  %s: N=%d, K=%d, worst rel %.3e (%.1f s with the oracles)" % (tag, sc.n, K_, worst,
                                                                        time.perf_counter() - t0), flush=True)
    return worst


def _solves_to_tolerance(torch, sc, solvers, tol=1e-5):
    """The step's right-hand side (advected, confined velocity's divergence) solved to tol by each
    solver through the device API; the pressure's relative residual is re-evaluated by the FP64
    oracle; each solve is repeated and must give the same bits."""
    topo = hns.create_index_grid(hns.GridIndexedData(sc.coords), sc.voxel_size)
    div = P.scene_divergence(sc, topo)
    o64 = Oracle(sc.coords, np.float64)
    d0 = _dense(torch, topo, div)
    out = {}
    for solver in solvers:
        p1, p2 = torch.empty_like(div), torch.empty_like(div)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = K.pressure_solve(topo, div, p1, sc.voxel_size, solver=solver, tol=tol, max_iters=20000)
        t1 = time.perf_counter()
        K.pressure_solve(topo, div, p2, sc.voxel_size, solver=solver, tol=tol, max_iters=20000)
        r2, d2 = o64.residual(_dense(torch, topo, p1), d0, sc.voxel_size)
        rel = float(np.sqrt(r2 / d2))
        same = bool(torch.equal(p1, p2))
        print("  %-5s %5d iterations, solver rel %.3e, FP64-oracle rel %.3e, %.1f ms, repeat bitwise equal %s" % (
            solver, st.iterations, st.rel_residual, rel, 1e3 * (t1 - t0), same), flush=True)
        out[solver] = (st, rel, same)
    return out


# ---- config 2 -------------------------------------------------------------------------------------

def test_cfg2_fire256_cone_steps_k100(hip):
    """Two fire steps on the literal cone (N = 1,976,832), K = 100, each against the FP32 oracle."""
    sc = _scene("fire256_cone")
    _case("cfg2 fire256_cone N=%d K=100, 2 steps" % sc.n)
    assert sc.n == 1976832
    t0 = time.perf_counter()
    worst = _per_step_parity("cone", sc.coords, sc.voxel_size, sc.velocity, sc.fields, sc.params, 100, sc.dt, 2)
    print("  worst rel %.3e (%.1f s)" % (worst, time.perf_counter() - t0))
    assert worst <= TOL_REL


def test_cfg2_fire256_8m_step_k100(hip):
    """One fire step on the 8M-voxel plume with the reference's 100 RBGS iterations."""
    sc = _scene("fire256_8m")
    _case("cfg2 fire256_8m N=%d K=100, 1 step" % sc.n)
    assert sc.n == 8355840
    assert _one_step_parity("8m", sc, 100) <= TOL_REL


def test_cfg2_fire256_8m_solves_to_tolerance(hip):
    torch = hip
    sc = _scene("fire256_8m")
    _case("cfg2 fire256_8m N=%d, MGCG and CG to 1e-5" % sc.n)
    res = _solves_to_tolerance(torch, sc, ("mgcg", "cg"))
    for solver, (st, rel, same) in res.items():
        assert rel <= 1.01e-5 and same, solver


def test_cfg2_fire256_8m_mgcg_step_is_deterministic(hip):
    """Two contexts from the same payload, one multigrid-to-1e-5 step each: every field bitwise equal."""
    sc = _scene("fire256_8m")
    _case("cfg2 fire256_8m N=%d, MGCG step determinism" % sc.n)
    data = _payload(sc.coords, sc.velocity, sc.fields)
    topo = hns.create_index_grid(data, sc.voxel_size)
    params = hns.make_params(**sc.params)
    outs = []
    for _ in range(2):
        ctx = hns.ensure_context(None, data, topo, sc.voxel_size)
        st = hns.advance(ctx, 0, sc.dt, params, solver="mgcg", tol=1e-5)
        out = _out_payload(sc.coords, ["density", "temperature", "fuel", "waste", "flame"])
        hns.download_for_output(ctx, out)
        outs.append(out)
        print("  mgcg step: %d iterations, rel %.3e" % (st.iterations, st.rel_residual), flush=True)
        assert st.rel_residual <= 1e-5
        hns.destroy_context(ctx)
    for k in outs[0].blocks:
        np.testing.assert_array_equal(outs[0][k], outs[1][k])


@pytest.mark.parametrize("name,calls", [("fire256_cone", 3), ("fire256_8m", 2)])
def test_cfg2_mgcg_steps_match_the_oracle(hip, name, calls):
    """The headline step: each call with the pressure solved by the multigrid-preconditioned CG to
    1e-5, against the oracle's step with its own multigrid-preconditioned CG to 1e-5 from the same
    state (oracle/hns_oracle_mg.cpp: the V-cycle bitwise the library's, the dot products summed in a
    different order): every field within TOL_REL and the same iteration count."""
    sc = _scene(name)
    _case("cfg2 %s N=%d, %d MGCG steps against the oracle's MGCG steps" % (name, sc.n, calls))
    data = _payload(sc.coords, sc.velocity, sc.fields)
    topo = hns.create_index_grid(data, sc.voxel_size)
    ctx = hns.ensure_context(None, data, topo, sc.voxel_size)
    params = hns.make_params(**sc.params)
    o32 = Oracle(sc.coords)
    vel, f = sc.velocity, dict(sc.fields)
    worst = 0.0
    for c in range(calls):
        st = hns.advance(ctx, 0, sc.dt, params, solver="mgcg", tol=1e-5)
        out = _out_payload(sc.coords, list(dict.fromkeys(list(f) + ["fuel", "waste", "temperature", "flame"])))
        hns.download_for_output(ctx, out)
        gv = out["velocity"].copy()
        gf = {k: out[k].copy() for k in out.blocks if k != "velocity" and ctx.has_field(k)}
        its = []
        ov, of = o32.advance(sc.voxel_size, 0, sc.dt, sc.params, vel, f, mgcg_tol=1e-5, solve_iterations=its)
        print("  call %d: gpu %d iterations (rel %.3e), oracle %s" % (c, st.iterations, st.rel_residual, its),
              flush=True)
        assert its and st.iterations == its[-1]
        worst = max(worst, _compare("%s#%d" % (name, c), gv, gf, ov, of))
        vel, f = gv, gf
    hns.destroy_context(ctx)
    assert worst <= TOL_REL


# ---- config 5 -------------------------------------------------------------------------------------

def test_cfg5_fire512_step_rbgs150(hip):
    """One large-fire step (cone x2 in 512^3) with the reference's 150 RBGS iterations."""
    sc = _scene("fire512")
    _case("cfg5 fire512 N=%d K=150, 1 step" % sc.n)
    assert sc.n > 12_000_000
    assert _one_step_parity("fire512", sc, 150) <= TOL_REL


def test_cfg5_fire512_solves_to_tolerance(hip):
    torch = hip
    sc = _scene("fire512")
    _case("cfg5 fire512 N=%d, MGCG and CG to 1e-5" % sc.n)
    res = _solves_to_tolerance(torch, sc, ("mgcg", "cg"))
    for solver, (st, rel, same) in res.items():
        assert rel <= 1.01e-5 and same, solver


def test_cfg5_fire512_adjoint_4_steps(hip):
    """A 4-step differentiable run (advect, 150 RBGS iterations, gradient subtraction, scalar
    advection per step) on fire512 with a checkpoint every 2 steps: recomputed segments reproduce
    the stored states bitwise, the device memory the pass holds is reported, and the gradient's
    directional derivative agrees with a central difference of the same forward (FP32)."""
    torch = hip
    sc = _scene("fire512")
    _case("cfg5 fire512 N=%d, 4-step checkpointed adjoint, K=150" % sc.n)
    topo = hns.create_index_grid(hns.GridIndexedData(sc.coords), sc.voxel_size)
    dx, dt, S, K_, every = sc.voxel_size, sc.dt, 4, 150, 2
    omega = float(np.float32(2.0 / (1.0 + np.sin(np.float32(3.14159) * np.float32(dx)))))
    u0 = torch.from_numpy(np.ascontiguousarray(sc.velocity.T)).cuda()
    # smooth densities (seeded Gaussian blobs, sigma 12 voxels) so the loss is smooth in the velocity
    # away from a few clamp switches; the scene's fuel is per-voxel noise, a kink at every cell
    blob = lambda seed: scenes.blobs(sc.coords, seed, 8, 12.0, (256, 160, 256), 100.0)  # noqa: E731
    rho0 = torch.from_numpy(blob(51)).cuda()
    target = torch.from_numpy(blob(52)).cuda()
    torch.cuda.synchronize()
    free0, total = torch.cuda.mem_get_info()
    vout, rout = torch.empty_like(u0), torch.empty_like(rho0)
    t0 = time.perf_counter()
    tape = diff.forward(topo, u0, rho0, S, dt, dx, K_, omega, every, vout, rout)
    torch.cuda.synchronize()
    free_fwd = torch.cuda.mem_get_info()[0]
    grho = (2.0 * (rout - target)).contiguous()
    gv0, gr0 = torch.empty_like(u0), torch.empty_like(rho0)
    tape.backward(None, grho, gv0, gr0)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    free_bwd = torch.cuda.mem_get_info()[0]
    info = tape.info()
    tape.close()
    held = max(free0 - free_fwd, free0 - free_bwd)
    print("  tape %s; device memory held by the pass (beyond inputs): after forward %.2f GB, after backward "
          "%.2f GB, peak of the two %.2f GB of %.0f GB; %.1f s" % (
              info, (free0 - free_fwd) / 1e9, (free0 - free_bwd) / 1e9, held / 1e9, total / 1e9, t1 - t0), flush=True)
    assert info["recompute_mismatches"] == 0 and info["checkpoints"] == 2
    assert torch.isfinite(gv0).all() and torch.isfinite(gr0).all()

    def loss(u, r):
        vo, ro = torch.empty_like(u), torch.empty_like(r)
        tp = diff.forward(topo, u, r, S, dt, dx, K_, omega, every, vo, ro)
        tp.close()
        return float(((ro.double() - target.double()) ** 2).sum())

    # density direction: for a fixed velocity the step is linear in rho (the MacCormack clamps aside,
    # which smooth blobs rarely reach), so the loss is quadratic in rho and the central difference is
    # exact at any step; FP32 states bound the agreement
    drho = torch.from_numpy(blob(53)).cuda()
    ad_r = float((gr0.double() * drho.double()).sum())
    fd_r = (loss(u0, rho0 + drho) - loss(u0, rho0 - drho)) / 2.0
    rel_r = abs(ad_r - fd_r) / abs(fd_r)
    print("  directional derivative along a smooth density direction: adjoint %.6e, central difference %.6e, "
          "rel %.2e" % (ad_r, fd_r, rel_r), flush=True)
    # velocity direction: every back-trace point moves, and trilinear sampling is only continuous across
    # cell faces, so the loss has kinks everywhere in aggregate and the difference quotient converges
    # only linearly in the step (a sanity check of the velocity adjoint at scale, not its precision)
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    du = torch.randn(u0.shape, device="cuda", generator=g) * 1e-2
    ad = float((gv0.double() * du.double()).sum())
    h = 0.25
    fd = (loss(u0 + h * du, rho0) - loss(u0 - h * du, rho0)) / (2 * h)
    rel_u = abs(ad - fd) / abs(fd)
    print("  directional derivative along a random velocity direction: adjoint %.6e, central difference "
          "(h = %.2f) %.6e, rel %.2e" % (ad, h, fd, rel_u), flush=True)
    assert rel_r <= 1e-3 and rel_u <= 3e-2


# ---- config 3 -------------------------------------------------------------------------------------

def test_cfg3_control128_adjoint(hip):
    """Differentiable density control at its stated size (control128: sphere R = 60 in 128^3,
    N = 1,177,088; 16 steps of advect + 80 RBGS iterations, checkpoints every 4). The GPU (FP32)
    gradient against FP64 central differences of the FP64 oracle at the 32 fixture entries
    (tests/golden/control128_fd.json, written by make_control128_fd.py): at every entry where the
    quotients at eps and eps/2 agree (no clamp switch or cell crossing within +-eps) it is within 1e-4
    of their Richardson extrapolation. The whole gradient is compared with the oracle's FP64 tape and
    with its FP32 tape (the same algorithm in FP32): the distance to the FP64 tape is the FP32
    algorithm's own, concentrated where FP32 and FP64 forwards take different clamp branches."""
    import json
    import os
    torch = hip
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "control128_fd.json")))
    sc = scenes.control128()
    _case("cfg3 control128 N=%d, %d steps, K=%d, checkpoint every 4" % (sc.n, fx["steps"], fx["iterations"]))
    assert sc.n == fx["voxels"] == 1177088 and fx["complete"]
    S, K_, dt, dx, om = fx["steps"], fx["iterations"], fx["dt"], fx["voxel_size"], fx["omega"]
    topo = hns.create_index_grid(hns.GridIndexedData(sc.coords), dx)
    u0 = torch.from_numpy(np.ascontiguousarray(sc.velocity.T)).cuda()
    rho0 = torch.from_numpy(sc.fields["density"]).cuda()
    tg = torch.from_numpy(sc.sources["target_density"]).cuda()
    vout, rout = torch.empty_like(u0), torch.empty_like(rho0)
    t0 = time.perf_counter()
    tape = diff.forward(topo, u0, rho0, S, dt, dx, K_, om, 4, vout, rout)
    loss = float(((rout.double() - tg.double()) ** 2).sum())
    gv0, gr0 = torch.empty_like(u0), torch.empty_like(rho0)
    tape.backward(None, (2.0 * (rout - tg)).contiguous(), gv0, gr0)
    torch.cuda.synchronize()
    info = tape.info()
    tape.close()
    print("  GPU forward + backward %.2f s; tape %s; loss gpu %.8e fp64 %.8e" % (
        time.perf_counter() - t0, info, loss, fx["loss_fp64"]), flush=True)
    assert info["recompute_mismatches"] == 0 and info["checkpoints"] == 4
    gv = gv0.cpu().numpy().reshape(3, -1).T.astype(np.float64)
    gr = gr0.cpu().numpy().astype(np.float64)
    gscale = fx["grad_scale_fp64"]
    rels, kept = [], 0
    for e in fx["entries"]:
        f1, f2 = e["fd_eps"], e["fd_half_eps"]
        g = gv[e["index"], e["component"]] if e["field"] == "velocity" else gr[e["index"]]
        smooth = abs(f1 - f2) <= 1e-5 * max(abs(f1), 1e-3 * gscale)
        fr = (4.0 * f2 - f1) / 3.0
        rel = abs(g - fr) / max(abs(fr), 1e-2 * gscale)
        print("  %-8s %7d/%d fd %.9e fd/2 %.9e -> %.9e  gpu %.9e  fp64 tape %.9e  rel %.2e%s" % (
            e["field"], e["index"], e["component"], f1, f2, fr, g, e["tape_fp64"], rel,
            "" if smooth else "  (kink inside +-eps: not compared)"))
        if smooth:
            kept += 1
            rels.append(rel)
    print("  %d of %d entries smooth; rel err vs FP64 FD: median %.2e, max %.2e" % (
        kept, len(fx["entries"]), np.median(rels), max(rels)), flush=True)
    assert kept >= 16 and max(rels) <= 1e-4
    # the whole gradient against the oracle's tapes (FP64, and the same algorithm in FP32)
    for dtype, bar in ((np.float64, 2e-3), (np.float32, 2e-3)):
        t1 = time.perf_counter()
        o = Oracle(sc.coords, dtype)
        _, tv, trr = o.diff_loss_grad(sc.velocity.astype(dtype), sc.fields["density"].astype(dtype),
                                      sc.sources["target_density"].astype(dtype), S, dt, dx, K_, om)
        ev = np.linalg.norm(gv - tv) / np.linalg.norm(tv)
        er = np.linalg.norm(gr - trr) / np.linalg.norm(trr)
        d2 = ((gv - tv) ** 2).ravel()
        top = np.sort(d2)[::-1][:max(1, d2.size // 10000)].sum() / max(d2.sum(), 1e-300)
        print("  vs the oracle's %s tape (%.0f s): rel L2 d/du0 %.2e, d/drho0 %.2e; share of the squared d/du0 "
              "difference in the top 0.01%% entries %.2f" % (np.dtype(dtype).name, time.perf_counter() - t1, ev, er,
                                                             top), flush=True)
        assert ev <= bar and er <= bar
