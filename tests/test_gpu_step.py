"""GPU parity of the whole step through the C ABI (HNS_* lifecycle and the standalone SOP entry
points) against the CPU oracle, plus the error paths and determinism of the boundary."""
import numpy as np
import pytest

import hnanosolver_amd as hns
from hnanosolver_amd import scenes
from oracle import Oracle

from _helpers import (TOL_REL, _compare, _payload, _per_step_parity, _run_gpu, _run_oracle, canonical_coords, max_abs,
                      plane_sdf, ragged_origins, ragged_scene, rel_err)

pytestmark = pytest.mark.gpu


def _case(name):
    print("CASE", name, flush=True)


def test_smoke64_steps(hip):
    """Three calls: the first advects only density (the fields needOrMake creates are not in the
    advNames snapshot, HNanoSolver.cu:248-250), the later ones all five scalars."""
    _case("smoke64 per-step x3")
    s = scenes.smoke64()
    assert _per_step_parity("smoke64", s.coords, s.voxel_size, s.velocity, s.fields, s.params, s.iterations, s.dt,
                            3) <= TOL_REL


def test_smoke64_multi_step_drift(hip):
    """Three calls chained on each side (no re-seeding): the GPU stays on the FP32 oracle."""
    _case("smoke64 chained x3")
    s = scenes.smoke64()
    args = (s.coords, s.voxel_size, s.velocity, s.fields, s.params, s.iterations, s.dt, 3)
    gv, gf = _run_gpu(*args)
    ov, of = _run_oracle(*args)
    o64 = _run_oracle(*args, dtype=np.float64)
    assert set(gf) == set(of)
    assert _compare("smoke64x3", gv, gf, ov, of, o64) <= TOL_REL


def test_fire_steps(hip):
    _case("fire precond128 (fire parameters, K=100) per-step x2")
    s = scenes.precond128(4)
    assert _per_step_parity("fire", s.coords, s.voxel_size, s.velocity, s.fields, s.params, s.iterations, s.dt,
                            2) <= TOL_REL


def test_all_features_step(hip):
    """Sources (src_vel, src_density, src_temperature, src_fuel), collision SDF, gravity,
    vorticity with factorScale 1, the disturbance re-solve and 2 substeps, on a ragged topology."""
    _case("all features ragged")
    c, u, f, dx, dt = ragged_scene(3, cfl=1.2)
    n = len(c)
    rng = np.random.default_rng(21)
    sources = {"src_density": rng.random(n).astype(np.float32), "src_temperature": rng.random(n).astype(np.float32) * 5,
               "src_fuel": rng.random(n).astype(np.float32) * 0.1}
    src_vel = (rng.standard_normal((n, 3)) * 0.5).astype(np.float32)
    sdf = plane_sdf(c, offset=-0.4, dx=dx)
    params = scenes.default_params(vorticityScale=0.3, factorScale=1.0, disturbanceEnable=1.0, substeps=2, gravity=2.0,
                                   ambientTemp=0.5)
    # collision branches (a back-trace landing inside the SDF) make this scene discontinuous in its
    # inputs: FP64 and FP32 runs of the same algorithm part ways, so parity is per step
    assert _per_step_parity("all", c, dx, u, f, params, 12, dt, 2, sources=sources, sdf=sdf,
                            src_vel=src_vel) <= TOL_REL


def test_standalone_entry_points(hip):
    _case("standalone AdvectIndexGrid / Velocity / Divergence / ProjectNonDivergent")
    c, u, f, dx, dt = ragged_scene(4, cfl=1.0)
    o = Oracle(c)
    d = hns.GridIndexedData(c)
    d.add_value_block("density", f["density"].copy())
    d.add_value_block("temperature", f["temperature"].copy())
    d.add_value_block("vel", u.copy())
    hns.advect_index_grid(d, dt, dx)
    for k in ("density", "temperature"):
        assert rel_err(d[k], o.advect_scalar(u, f[k], dt, 1.0 / dx)) <= TOL_REL
    dv = hns.GridIndexedData(c)
    dv.add_value_block("vel", u.copy())
    hns.advect_index_grid_velocity(dv, dt, dx)
    assert rel_err(dv["vel"], o.advect_vector(u, dt, 1.0 / dx)) <= TOL_REL
    dd = hns.GridIndexedData(c)
    dd.add_value_block("vel", u.copy())
    dd.add_value_block("divergence", np.zeros(len(c), np.float32))
    hns.divergence(dd, dx)
    div = o.divergence(u, 1.0 / dx)
    assert rel_err(dd["divergence"], div) <= TOL_REL
    dp = hns.GridIndexedData(c)
    dp.add_value_block("vel", u.copy())
    K = 20
    hns.project_non_divergent(dp, K, dx)
    omega = np.float32(2.0 / (1.0 + np.sin(3.14159 * np.float64(np.float32(dx)))))
    p = np.zeros(len(c), np.float32)
    for _ in range(K):
        p = o.rbgs(div, p, dx, 0, omega)
        p = o.rbgs(div, p, dx, 1, omega)
    ref = o.subtract_gradient(u, p, 1.0 / dx)
    e = rel_err(dp["vel"], ref)
    print("  project rel %.3e max-abs %.3e" % (e, max_abs(dp["vel"], ref)))
    assert e <= TOL_REL


def test_context_reuse_and_recreate(hip):
    """HNS_EnsureContext keeps device state when N and voxel size are unchanged and re-seeds
    from the host payload otherwise (HNanoSolver.cu:64-155)."""
    _case("context reuse")
    s = scenes.smoke64()
    data = _payload(s.coords, s.velocity, s.fields)
    topo = hns.create_index_grid(data, s.voxel_size)
    ctx = hns.ensure_context(None, data, topo, s.voxel_size)
    p = hns.make_params(**s.params)
    hns.advance(ctx, 5, s.dt, p)
    h0 = ctx.handle.value
    ctx = hns.ensure_context(ctx, data, topo, s.voxel_size)  # reuse: state kept
    assert ctx.handle.value == h0
    assert ctx.has_field("fuel")
    ctx = hns.ensure_context(ctx, data, topo, s.voxel_size * 2)  # voxel size changed: new context
    assert not ctx.has_field("fuel") and ctx.has_field("density")
    n, L, ns = ctx.info()
    assert (n, L, ns) == (s.n, s.n_leaves, 1)
    m = hns.get_active_mask(ctx)
    assert m.shape == (s.n,) and (m == 1).all()
    hns.destroy_context(ctx)
    assert not ctx
    assert hns.get_active_mask(None).size == 0


def test_boundary_errors(hip):
    _case("boundary errors")
    c = canonical_coords(ragged_origins(5))
    # leaves out of voxelsToGrid order
    L = len(c) // 512
    swapped = c.reshape(L, 512, 3)[[1, 0] + list(range(2, L))].reshape(-1, 3)
    with pytest.raises(hns.HNSError) as e:
        hns.create_index_grid(hns.GridIndexedData(swapped), 0.1)
    assert e.value.status == -5
    # not leaf-dense (voxel order inside a leaf broken)
    bad = c.copy()
    bad[[3, 4]] = bad[[4, 3]]
    with pytest.raises(hns.HNSError) as e:
        hns.create_index_grid(hns.GridIndexedData(bad), 0.1)
    assert e.value.status == -4
    with pytest.raises(hns.HNSError) as e:
        hns.create_index_grid(hns.GridIndexedData(c[:700]), 0.1)
    assert e.value.status == -4
    with pytest.raises(hns.HNSError) as e:
        hns.create_index_grid(hns.GridIndexedData(np.zeros((0, 3), np.int32)), 0.1)
    assert e.value.status == -2
    # no velocity block
    d = hns.GridIndexedData(c)
    d.add_value_block("density", np.zeros(len(c), np.float32))
    topo = hns.create_index_grid(d, 0.1)
    with pytest.raises(hns.HNSError) as e:
        hns.ensure_context(None, d, topo, 0.1)
    assert e.value.status == -6
    # payload / topology size mismatch
    d2 = hns.GridIndexedData(c[:512])
    d2.add_value_block("velocity", np.zeros((512, 3), np.float32))
    with pytest.raises(hns.HNSError) as e:
        hns.ensure_context(None, d2, topo, 0.1)
    assert e.value.status == -4


def test_determinism(hip):
    _case("determinism fire")
    s = scenes.precond128(6)
    args = (s.coords, s.voxel_size, s.velocity, s.fields, s.params, 30, s.dt, 2)
    a = _run_gpu(*args)
    b = _run_gpu(*args)
    np.testing.assert_array_equal(a[0], b[0])
    for k in a[1]:
        np.testing.assert_array_equal(a[1][k], b[1][k])


def test_rbgs_to_tolerance(hip):
    _case("rbgs to tolerance smoke64")
    s = scenes.smoke64()
    data = _payload(s.coords, s.velocity, s.fields)
    topo = hns.create_index_grid(data, s.voxel_size)
    ctx = hns.ensure_context(None, data, topo, s.voxel_size)
    p = hns.make_params(**s.params)
    st = hns.advance(ctx, 0, s.dt, p, solver="rbgs_tol", tol=1e-3, check_every=25, max_iters=20000)
    print("  iterations %d rel residual %.3e" % (st.iterations, st.rel_residual))
    assert st.rel_residual <= 1e-3 and st.iterations > 0
    st2 = hns.advance(ctx, 60, s.dt, p, stats=True)
    assert st2.iterations == 60 and 0 < st2.rel_residual < 1
    hns.destroy_context(ctx)
