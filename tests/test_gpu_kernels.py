"""GPU parity of every kernel of the step against the CPU oracle (same seeded inputs).

Each test prints the case it runs (scene, kernel) before touching the GPU. Tolerance: the FP32
bar of the step, max|gpu - oracle| <= 1e-5 * max|oracle|; max-abs errors are printed."""
import numpy as np
import pytest

import hnanosolver_amd as hns
from hnanosolver_amd import kernels as K
from hnanosolver_amd import scenes
from oracle import Oracle

from _helpers import TOL_REL, canonical_coords, max_abs, plane_sdf, ragged_scene, rel_err

pytestmark = pytest.mark.gpu


def _case(name):
    print("CASE", name, flush=True)


def _scene(name):
    if name == "smoke64":
        s = scenes.smoke64()
        return s.coords, s.velocity, s.fields["density"], s.voxel_size, s.dt
    if name == "ragged":
        c, u, f, dx, dt = ragged_scene(0, cfl=1.5)
        return c, u, f["density"], dx, dt
    if name == "ragged_fast":  # CFL ~ 11 voxels: back-traces leave the 27-leaf neighbourhood
        c, u, f, dx, dt = ragged_scene(1, cfl=11.0)
        return c, u, f["density"], dx, dt
    raise KeyError(name)


def _topo(coords, dx):
    d = hns.GridIndexedData(coords)
    return hns.create_index_grid(d, dx)


def _soa(torch, vel):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(vel, np.float32).T)).cuda()


def _aos(t):
    return t.cpu().numpy().T.copy()


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def _report(tag, got, ref):
    e, m = rel_err(got, ref), max_abs(got, ref)
    diff = int(np.count_nonzero(np.asarray(got, np.float32) != np.asarray(ref, np.float32)))
    print("  %-28s rel %.3e  max-abs %.3e  elements differing from the oracle: %d of %d" % (
        tag, e, m, diff, np.size(ref)), flush=True)
    return e


SCENES = ["smoke64", "ragged", "ragged_fast"]


@pytest.mark.parametrize("name", SCENES)
def test_topology_neighbour_rows(hip, name):
    _case("topology " + name)
    c, *_ = _scene(name)
    t = _topo(c, 0.1)
    nb = t.neighbours()
    o = Oracle(c)
    org = c.reshape(-1, 512, 3)[:, 0, :]
    L = len(org)
    assert nb.shape == (L, 27)
    k = np.arange(27)
    d = np.stack([k // 9 - 1, (k // 3) % 3 - 1, k % 3 - 1], -1) * 8
    exp = np.empty((L, 27), np.int64)
    for l in range(L):
        for q in range(27):
            off = o.offset(*(org[l] + d[q]))
            exp[l, q] = -1 if off == 0 else (off - 1) // 512
    np.testing.assert_array_equal(nb, exp)


@pytest.mark.parametrize("name", SCENES)
def test_advect_velocity(hip, name):
    torch = hip
    _case("advect_velocity " + name)
    c, u, _, dx, dt = _scene(name)
    t = _topo(c, dx)
    o = Oracle(c)
    ref = o.advect_vector(u, dt, 1.0 / dx)
    vin = _soa(torch, u)
    out = torch.empty_like(vin)
    K.advect_velocity(t, vin, out, dt, 1.0 / dx)
    assert _report("advect_velocity", _aos(out), ref) <= TOL_REL


@pytest.mark.parametrize("name", SCENES)
def test_advect_scalars(hip, name):
    torch = hip
    _case("advect_scalars " + name)
    c, u, dens, dx, dt = _scene(name)
    t = _topo(c, dx)
    o = Oracle(c)
    rng = np.random.default_rng(3)
    fields = [dens, rng.random(len(c)).astype(np.float32), (dens * dens).astype(np.float32)]
    refs = o.advect_scalars(u, fields, dt, 1.0 / dx)
    vin = _soa(torch, u)
    ins = [_dev(torch, f) for f in fields]
    outs = [torch.empty_like(x) for x in ins]
    K.advect_scalars(t, vin, ins, outs, dt, 1.0 / dx)
    for i in range(3):
        assert _report("advect_scalars[%d]" % i, outs[i].cpu().numpy(), refs[i]) <= TOL_REL


def test_advect_scalars_nine_fields(hip):
    """More fields than one launch carries (8): the host splits them over launches."""
    torch = hip
    _case("advect_scalars 9 fields")
    c, u, dens, dx, dt = _scene("ragged")
    t = _topo(c, dx)
    o = Oracle(c)
    fields = [(dens * (k + 1)).astype(np.float32) for k in range(9)]
    refs = o.advect_scalars(u, fields, dt, 1.0 / dx)
    vin = _soa(torch, u)
    ins = [_dev(torch, f) for f in fields]
    outs = [torch.empty_like(x) for x in ins]
    K.advect_scalars(t, vin, ins, outs, dt, 1.0 / dx)
    for i in range(9):
        assert rel_err(outs[i].cpu().numpy(), refs[i]) <= TOL_REL


@pytest.mark.parametrize("name", SCENES)
def test_advect_scalar_single(hip, name):
    torch = hip
    _case("advect_scalar " + name)
    c, u, dens, dx, dt = _scene(name)
    t = _topo(c, dx)
    ref = Oracle(c).advect_scalar(u, dens, dt, 1.0 / dx)
    out = torch.empty(len(c), device="cuda")
    K.advect_scalar(t, _soa(torch, u), _dev(torch, dens), out, dt, 1.0 / dx)
    assert _report("advect_scalar", out.cpu().numpy(), ref) <= TOL_REL


@pytest.mark.parametrize("name,F", [("smoke64", 1.0), ("ragged", 1.0), ("ragged", 2.7), ("ragged", -1.0)])
def test_vorticity_forces(hip, name, F):
    torch = hip
    _case("vorticity %s F=%g" % (name, F))
    c, u, _, dx, dt = _scene(name)
    t = _topo(c, dx)
    o = Oracle(c)
    ref = o.vorticity(u, dt, 1.0 / dx, 0.2, F)
    rng = np.random.default_rng(5)
    src = rng.standard_normal((len(c), 3)).astype(np.float32)
    g = 9.8
    ref2 = ref + src * np.float32(dt)
    ref2[:, 1] -= np.float32(g * dt)
    vin = _soa(torch, u)
    out = torch.empty_like(vin)
    K.vorticity_forces(t, vin, out, dt, 1.0 / dx, 0.2, F)
    assert _report("vorticity", _aos(out), ref) <= TOL_REL
    K.vorticity_forces(t, vin, out, dt, 1.0 / dx, 0.2, F, src_vel=_soa(torch, src), gravity=g)
    assert _report("vorticity+src+gravity", _aos(out), ref2) <= TOL_REL


@pytest.mark.parametrize("name", SCENES)
def test_divergence(hip, name):
    torch = hip
    _case("divergence " + name)
    c, u, _, dx, dt = _scene(name)
    t = _topo(c, dx)
    ref = Oracle(c).divergence(u, 1.0 / dx)
    vin = _soa(torch, u)
    d = torch.empty(len(c), device="cuda")
    K.divergence(t, vin, d, 1.0 / dx)
    assert _report("divergence", d.cpu().numpy(), ref) <= TOL_REL
    dcs = torch.empty_like(d)
    K.divergence(t, vin, dcs, 1.0 / dx, colour_split=True)
    back = torch.empty_like(d)
    K.from_colour_split(t, dcs, back)
    np.testing.assert_array_equal(back.cpu().numpy(), d.cpu().numpy())


@pytest.mark.parametrize("name", SCENES)
def test_rbgs_sweeps_and_residual(hip, name):
    torch = hip
    _case("rbgs " + name)
    c, u, _, dx, dt = _scene(name)
    t = _topo(c, dx)
    o = Oracle(c)
    rng = np.random.default_rng(7)
    div = rng.standard_normal(len(c)).astype(np.float32)
    p = rng.standard_normal(len(c)).astype(np.float32) * 0.01
    omega = np.float32(2.0 / (1.0 + np.sin(np.float32(3.14159) * np.float32(dx))))
    ref = p.copy()
    for it in range(4):
        ref = o.rbgs(div, ref, dx, 0, omega)
        ref = o.rbgs(div, ref, dx, 1, omega)
    dcs = torch.empty(len(c), device="cuda")
    pcs = torch.empty(len(c), device="cuda")
    K.to_colour_split(t, _dev(torch, div), dcs)
    K.to_colour_split(t, _dev(torch, p), pcs)
    for it in range(4):
        K.rbgs_sweep(t, dcs, pcs, dx, 0, float(omega))
        K.rbgs_sweep(t, dcs, pcs, dx, 1, float(omega))
    got = torch.empty_like(pcs)
    K.from_colour_split(t, pcs, got)
    assert _report("rbgs x4", got.cpu().numpy(), ref) <= TOL_REL
    r2, d2 = K.residual(t, pcs, dcs, dx)
    rr2, rd2 = o.residual(ref, div, dx)
    print("  residual gpu %.6e oracle %.6e" % (np.sqrt(r2 / d2), np.sqrt(rr2 / rd2)))
    assert abs(d2 - rd2) <= 1e-9 * rd2
    assert abs(np.sqrt(r2) - np.sqrt(rr2)) <= 1e-4 * np.sqrt(rr2)


@pytest.mark.parametrize("name", SCENES)
def test_subtract_gradient(hip, name):
    torch = hip
    _case("subtract_gradient " + name)
    c, u, _, dx, dt = _scene(name)
    t = _topo(c, dx)
    rng = np.random.default_rng(9)
    p = rng.standard_normal(len(c)).astype(np.float32)
    ref = Oracle(c).subtract_gradient(u, p, 1.0 / dx)
    pcs = torch.empty(len(c), device="cuda")
    K.to_colour_split(t, _dev(torch, p), pcs)
    vin = _soa(torch, u)
    out = torch.empty_like(vin)
    K.subtract_gradient(t, vin, pcs, out, 1.0 / dx)
    assert _report("subtract_gradient", _aos(out), ref) <= TOL_REL


def test_collision_paths(hip):
    """SDF branches of advect_vector, advect_scalars, subtractPressureGradient and
    enforceCollisionBoundaries (Kernel.cu:132-158, :195-248, :374-423, :528-538)."""
    torch = hip
    _case("collision ragged")
    c, u, dens, dx, dt = _scene("ragged")
    t = _topo(c, dx)
    o = Oracle(c)
    sdf = plane_sdf(c, offset=0.05, dx=dx)
    assert (sdf < 0).any() and ((sdf >= 0) & (sdf < 0.1)).any()
    vin = _soa(torch, u)
    dsdf = _dev(torch, sdf)
    out = torch.empty_like(vin)
    K.advect_velocity(t, vin, out, dt, 1.0 / dx, sdf=dsdf)
    assert _report("advect_velocity+sdf", _aos(out), o.advect_vector(u, dt, 1.0 / dx, sdf)) <= TOL_REL
    ins = [_dev(torch, dens)]
    outs = [torch.empty_like(ins[0])]
    K.advect_scalars(t, vin, ins, outs, dt, 1.0 / dx, sdf=dsdf)
    assert _report("advect_scalars+sdf", outs[0].cpu().numpy(), o.advect_scalars(u, [dens], dt, 1.0 / dx, sdf)[0]) <= TOL_REL
    p = np.random.default_rng(1).standard_normal(len(c)).astype(np.float32)
    pcs = torch.empty(len(c), device="cuda")
    K.to_colour_split(t, _dev(torch, p), pcs)
    K.subtract_gradient(t, vin, pcs, out, 1.0 / dx, sdf=dsdf)
    assert _report("gradient+sdf", _aos(out), o.subtract_gradient(u, p, 1.0 / dx, sdf)) <= TOL_REL
    v2 = vin.clone()
    K.enforce_collision(t, v2, dsdf, dx)
    assert _report("enforce_collision", _aos(v2), o.enforce_collision(u, sdf, dx)) <= TOL_REL


def test_combustion_buoyancy(hip):
    torch = hip
    _case("combustion_buoyancy")
    c, u, f, dx, dt = ragged_scene(2)
    n = len(c)
    rng = np.random.default_rng(11)
    fuel = rng.random(n).astype(np.float32)
    fuel[::7] = 0.0005  # snaps to 0
    waste = (rng.random(n) * 0.9).astype(np.float32)  # some oxygen < 0
    temp = (rng.random(n) * 40).astype(np.float32)
    flame = rng.random(n).astype(np.float32) * 0.2
    div = rng.standard_normal(n).astype(np.float32)
    t = _topo(c, dx)
    o = Oracle(c)
    rf, rw, rt, rd, rfl = o.combustion(fuel, waste, temp, div, flame, 0.9, 0.1)
    rv = o.buoyancy(u, rt, dt, 23.0, 1.0)
    d = [_dev(torch, a) for a in (fuel, waste, temp, flame, div)]
    vin = _soa(torch, u)
    vout = torch.empty_like(vin)
    K.combustion_buoyancy(t, d[0], d[1], d[2], d[3], d[4], vin, vout, dt, 0.9, 0.1, 23.0, 1.0)
    for tag, got, ref in (("fuel", d[0], rf), ("waste", d[1], rw), ("temp", d[2], rt), ("flame", d[3], rfl),
                          ("div", d[4], rd)):
        assert _report(tag, got.cpu().numpy(), ref) <= TOL_REL
    assert _report("buoyancy", _aos(vout), rv) <= TOL_REL


def test_layout_roundtrips(hip):
    torch = hip
    _case("layouts")
    c, u, dens, dx, dt = _scene("ragged")
    t = _topo(c, dx)
    x = _dev(torch, dens)
    cs = torch.empty_like(x)
    back = torch.empty_like(x)
    K.to_colour_split(t, x, cs)
    K.from_colour_split(t, cs, back)
    assert torch.equal(back, x)
    aos = _dev(torch, u.reshape(-1))
    soa = torch.empty_like(aos)
    aos2 = torch.empty_like(aos)
    K.aos_to_soa(aos, soa)
    K.soa_to_aos(soa, aos2)
    assert torch.equal(aos2, aos)
    np.testing.assert_array_equal(soa.cpu().numpy().reshape(3, -1).T, u)


def test_cg_against_scipy_fixture_and_oracle(hip):
    """GPU CG (colour-split FP32 vectors, FP64 dots) on the committed scipy direct solve, and on
    the divergence of the first smoke64 step against the FP64 oracle CG."""
    import os
    torch = hip
    _case("cg poisson_block + smoke64")
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "poisson_block.npz"))
    c, div, p_ref, dx = fx["coords"], fx["div"], fx["p"], float(fx["dx"])
    t = _topo(c, dx)
    dcs = torch.empty(len(c), device="cuda")
    pcs = torch.empty(len(c), device="cuda")
    K.to_colour_split(t, _dev(torch, div), dcs)
    st = K.pressure_solve(t, dcs, pcs, dx, solver="cg", tol=1e-5)
    got = torch.empty_like(pcs)
    K.from_colour_split(t, pcs, got)
    p = got.cpu().numpy()
    o = Oracle(c, np.float64)
    r2, d2 = o.residual(p, div, dx)
    print("  block: iterations %d, gpu rel %.3e, oracle-checked rel %.3e, |p - p_scipy| %.3e" % (
        st.iterations, st.rel_residual, np.sqrt(r2 / d2), np.abs(p - p_ref).max()))
    assert st.rel_residual <= 1e-5 and np.sqrt(r2 / d2) <= 1.01e-5
    assert np.abs(p - p_ref).max() <= 1e-4 * np.abs(p_ref).max()
    s = scenes.smoke64()
    o32 = Oracle(s.coords)
    d0 = o32.divergence(o32.advect_vector(s.velocity, s.dt, 1 / s.voxel_size), 1 / s.voxel_size)
    t = _topo(s.coords, s.voxel_size)
    dcs = torch.empty(s.n, device="cuda")
    pcs = torch.empty(s.n, device="cuda")
    K.to_colour_split(t, _dev(torch, d0), dcs)
    st = K.pressure_solve(t, dcs, pcs, s.voxel_size, solver="cg", tol=1e-5)
    _, it64, _ = Oracle(s.coords, np.float64).cg(d0, s.voxel_size, tol=1e-5)
    got = torch.empty_like(pcs)
    K.from_colour_split(t, pcs, got)
    r2, d2 = Oracle(s.coords, np.float64).residual(got.cpu().numpy(), d0, s.voxel_size)
    print("  smoke64: gpu iterations %d (FP64 oracle %d), rel %.3e (oracle-checked %.3e)" % (
        st.iterations, it64, st.rel_residual, np.sqrt(r2 / d2)))
    assert np.sqrt(r2 / d2) <= 1.01e-5
    assert abs(st.iterations - it64) <= max(3, 0.15 * it64)


def test_rbgs_to_tolerance_device_api(hip):
    torch = hip
    _case("rbgs_tol poisson_block")
    import os
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "poisson_block.npz"))
    c, div, p_ref, dx = fx["coords"], fx["div"], fx["p"], float(fx["dx"])
    t = _topo(c, dx)
    dcs = torch.empty(len(c), device="cuda")
    pcs = torch.empty(len(c), device="cuda")
    K.to_colour_split(t, _dev(torch, div), dcs)
    omega = float(np.float32(2.0 / (1.0 + np.sin(np.float32(3.14159) * np.float32(dx)))))
    st = K.pressure_solve(t, dcs, pcs, dx, omega=omega, solver="rbgs_tol", tol=1e-5, check_every=10)
    got = torch.empty_like(pcs)
    K.from_colour_split(t, pcs, got)
    r2, d2 = Oracle(c, np.float64).residual(got.cpu().numpy(), div, dx)
    print("  iterations %d rel %.3e (oracle-checked %.3e)" % (st.iterations, st.rel_residual, np.sqrt(r2 / d2)))
    assert st.rel_residual <= 1e-5 and np.sqrt(r2 / d2) <= 1.01e-5


@pytest.mark.parametrize("name,iters", [("ragged", 3), ("ragged", 4), ("smoke64", 7)])
def test_blocked_iteration_matches_colour_sweeps_bitwise(hip, name, iters):
    """The LDS temporally blocked full iteration (used by the solver) and the per-colour sweep
    kernel compute the same bits from p = 0."""
    torch = hip
    _case("rbgs blocked vs sweeps %s K=%d" % (name, iters))
    c, u, _, dx, dt = _scene(name)
    t = _topo(c, dx)
    rng = np.random.default_rng(13)
    div = rng.standard_normal(len(c)).astype(np.float32)
    omega = float(np.float32(2.0 / (1.0 + np.sin(np.float32(3.14159) * np.float32(dx)))))
    dcs = torch.empty(len(c), device="cuda")
    K.to_colour_split(t, _dev(torch, div), dcs)
    p_blocked = torch.full((len(c),), float("nan"), device="cuda")
    import os
    os.environ["HNS_RBGS_KERNEL"] = "blocked"
    try:
        K.pressure_solve(t, dcs, p_blocked, dx, omega=omega, iterations=iters, solver="rbgs")
    finally:
        del os.environ["HNS_RBGS_KERNEL"]
    p_sweep = torch.zeros(len(c), device="cuda")
    for _ in range(iters):
        K.rbgs_sweep(t, dcs, p_sweep, dx, 0, omega)
        K.rbgs_sweep(t, dcs, p_sweep, dx, 1, omega)
    assert torch.equal(p_blocked, p_sweep)
