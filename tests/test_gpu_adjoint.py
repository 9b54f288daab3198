"""GPU reverse mode of the differentiable step against the FP64 host tape and FP64 finite
differences.

Per kernel: the transposed RBGS iterations are deterministic and compared with the FP32 oracle's
transpose (max|gpu - oracle| <= 1e-6 max|oracle|); the advection transposes scatter with float
atomics (order not fixed), compared with the FP32 oracle to 1e-5 max|oracle|. The whole
checkpointed adjoint (FP32 on the GPU) is compared with the FP64 host tape (relative L2 error) and
with FP64 central differences (eps 1e-3) at 32 random entries; recomputed checkpoint segments must
reproduce the stored states bitwise."""
import numpy as np
import pytest

import hnanosolver_amd as hns
from hnanosolver_amd import diff
from hnanosolver_amd import kernels as K
from hnanosolver_amd import scenes
from oracle import Oracle

from _helpers import canonical_coords, ragged_origins

pytestmark = pytest.mark.gpu


def _case(name):
    print("CASE", name, flush=True)


def _topo(coords, dx):
    return hns.create_index_grid(hns.GridIndexedData(coords), dx)


def _omega(dx):
    return float(np.float32(2.0 / (1.0 + np.sin(np.float32(3.14159) * np.float32(dx)))))


def _soa(torch, v):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(v, np.float32).T)).cuda()


def _aos(t):
    return t.cpu().numpy().reshape(3, -1).T.astype(np.float64)


def small_scene(seed=0, cfl=0.8):
    c = canonical_coords(ragged_origins(seed)[:40])
    x = c.astype(np.float64)
    rng = np.random.default_rng(seed)
    u = np.stack([np.sin(0.2 * x[:, 1] + 0.3), np.cos(0.17 * x[:, 0]), 0.5 * np.sin(0.1 * x[:, 0] + 0.2 * x[:, 1])], -1)
    u += 0.2 * rng.standard_normal(u.shape)
    dx, dt = 0.1, 0.05
    u *= cfl * dx / dt / np.abs(u).max()
    rho = np.exp(-((x - x.mean(0)) ** 2).sum(1) / (2 * 6.0 ** 2)) + 0.1 * np.sin(0.3 * x[:, 2])
    return c, u.astype(np.float32), rho.astype(np.float32), dx, dt


@pytest.mark.parametrize("K_", [1, 3, 20])
def test_rbgs_transpose_matches_oracle(hip, K_):
    torch = hip
    s = scenes.smoke64()
    _case("rbgs transpose K=%d smoke64" % K_)
    t = _topo(s.coords, s.voxel_size)
    g = np.random.default_rng(K_).standard_normal(s.n).astype(np.float32)
    ref = Oracle(s.coords).rbgs_adj(g, s.voxel_size, K_, _omega(s.voxel_size))
    gd = torch.from_numpy(g).cuda()
    g_cs = torch.empty_like(gd)
    K.to_colour_split(t, gd, g_cs)
    gdiv_cs = torch.zeros_like(gd)
    diff.rbgs_adj(t, g_cs, gdiv_cs, s.voxel_size, K_, _omega(s.voxel_size))
    out = torch.empty_like(gd)
    K.from_colour_split(t, gdiv_cs, out)
    got = out.cpu().numpy()
    err = np.abs(got - ref).max()
    print("  max|d| %.2e of max %.2e, differing %d of %d" % (err, np.abs(ref).max(), np.count_nonzero(got != ref), s.n))
    assert err <= 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("cfl", [0.8, 3.5])
def test_advection_transposes_match_oracle(hip, cfl):
    """cfl 3.5 sends traces beyond the 2-voxel LDS halo, through the far-target path."""
    torch = hip
    c, u, rho, dx, dt = small_scene(7, cfl=cfl)
    _case("advection transposes, ragged scene, CFL %.1f" % cfl)
    t = _topo(c, dx)
    o = Oracle(c)
    rng = np.random.default_rng(8)
    gout = rng.standard_normal(len(c)).astype(np.float32)
    Gout = rng.standard_normal((len(c), 3)).astype(np.float32)
    inv_dx = float(np.float32(1.0) / np.float32(dx))
    gin_ref, gvel_ref = o.advect_scalars_adj(u, rho, gout, dt, inv_dx)
    gv2_ref = o.advect_vector_adj(u, Gout, dt, inv_dx)
    ud = _soa(torch, u)
    gin = torch.zeros(len(c), device="cuda")
    gvel = torch.zeros_like(ud)
    diff.advect_scalars_adj(t, ud, torch.from_numpy(rho).cuda(), torch.from_numpy(gout).cuda(), gin, gvel, dt, inv_dx)
    gv2 = torch.zeros_like(ud)
    diff.advect_vector_adj(t, ud, _soa(torch, Gout), gv2, dt, inv_dx)
    for name, got, ref in (("scalars d/dfield", gin.cpu().numpy(), gin_ref), ("scalars d/dvel", _aos(gvel), gvel_ref),
                           ("vector d/dvel", _aos(gv2), gv2_ref)):
        err = np.abs(got - ref).max()
        print("  %-18s max|d| %.2e of max %.2e" % (name, err, np.abs(ref).max()))
        assert err <= 1e-5 * np.abs(ref).max()


def _loss_grad_gpu(torch, t, u, rho, target, steps, dt, dx, K_, every):
    ud, rd = _soa(torch, u), torch.from_numpy(np.asarray(rho, np.float32)).cuda()
    vout, rout = torch.empty_like(ud), torch.empty_like(rd)
    tape = diff.forward(t, ud, rd, steps, dt, dx, K_, _omega(dx), every, vout, rout)
    tg = torch.from_numpy(np.asarray(target, np.float32)).cuda()
    loss = float(((rout.double() - tg.double()) ** 2).sum())
    grho = (2.0 * (rout - tg)).contiguous()
    gv0, gr0 = torch.empty_like(ud), torch.empty_like(rd)
    tape.backward(None, grho, gv0, gr0)
    info = tape.info()
    tape.close()
    return loss, _aos(gv0), gr0.cpu().numpy().astype(np.float64), info


def test_checkpointed_adjoint_matches_fp64_tape(hip):
    """The density-control problem at test size (control64: 16 steps, 80 RBGS, checkpoints every 4)."""
    torch = hip
    sc = scenes.control64()
    target = sc.sources["target_density"]
    rho = sc.fields["density"]
    S, K_, every = 16, 80, 4
    _case("checkpointed adjoint control64 N=%d S=%d K=%d every=%d" % (sc.n, S, K_, every))
    t = _topo(sc.coords, sc.voxel_size)
    loss, gv, gr, info = _loss_grad_gpu(torch, t, sc.velocity, rho, target, S, sc.dt, sc.voxel_size, K_, every)
    print("  tape %s" % info)
    assert info["recompute_mismatches"] == 0 and info["checkpoints"] == 4
    o64 = Oracle(sc.coords, np.float64)
    loss64, gv64, gr64 = o64.diff_loss_grad(sc.velocity.astype(np.float64), rho.astype(np.float64),
                                            target.astype(np.float64), S, sc.dt, sc.voxel_size, K_,
                                            _omega(sc.voxel_size))
    ev = np.linalg.norm(gv - gv64) / np.linalg.norm(gv64)
    er = np.linalg.norm(gr - gr64) / np.linalg.norm(gr64)
    print("  loss gpu %.6e fp64 %.6e; rel L2 err d/du0 %.2e, d/drho0 %.2e" % (loss, loss64, ev, er))
    assert abs(loss - loss64) <= 1e-4 * loss64
    assert ev <= 2e-3 and er <= 2e-3
    # every step stored: the same gradients up to the order of the float atomics
    _, gv1, gr1, info1 = _loss_grad_gpu(torch, t, sc.velocity, rho, target, S, sc.dt, sc.voxel_size, K_, 1)
    assert info1["checkpoints"] == 16
    assert np.linalg.norm(gv1 - gv) <= 1e-5 * np.linalg.norm(gv)
    assert np.linalg.norm(gr1 - gr) <= 1e-5 * np.linalg.norm(gr)


def test_adjoint_matches_fp64_central_differences(hip):
    """GPU (FP32) gradient against FP64 central differences (eps 1e-3) of the FP64 forward at 32
    random entries of (u0, rho0) where the loss is smooth across +-eps."""
    torch = hip
    c, u, rho, dx, dt = small_scene(5, cfl=0.6)
    _case("adjoint vs FP64 FD, ragged scene N=%d, 4 steps, K=20" % len(c))
    target = (np.roll(rho, 37) * 0.8).astype(np.float32)
    S, K_ = 4, 20
    t = _topo(c, dx)
    _, gv, gr, info = _loss_grad_gpu(torch, t, u, rho, target, S, dt, dx, K_, 2)
    assert info["recompute_mismatches"] == 0
    o = Oracle(c, np.float64)
    u64, r64, tg64 = u.astype(np.float64), rho.astype(np.float64), target.astype(np.float64)

    def L(uu, rr):
        _, r = o.diff_forward(uu, rr, S, dt, dx, K_, _omega(dx))
        return float(((r - tg64) ** 2).sum())

    rng = np.random.default_rng(9)
    cand_v = np.argsort(-np.abs(gv).max(1))[:300]
    cand_r = np.argsort(-np.abs(gr))[:300]
    eps, rels, kept = 1e-3, [], 0
    gscale = max(np.abs(gv).max(), np.abs(gr).max())
    for j in range(32):
        if j % 2 == 0:
            i = int(rng.choice(cand_v))
            k = int(rng.integers(3))
            g = gv[i, k]

            def fd(e):
                up, um = u64.copy(), u64.copy()
                up[i, k] += e
                um[i, k] -= e
                return (L(up, r64) - L(um, r64)) / (2 * e)
        else:
            i = int(rng.choice(cand_r))
            g = gr[i]

            def fd(e):
                rp, rm = r64.copy(), r64.copy()
                rp[i] += e
                rm[i] -= e
                return (L(u64, rp) - L(u64, rm)) / (2 * e)
        f1, f2 = fd(eps), fd(eps / 2)
        if abs(f1 - f2) > 1e-5 * max(abs(f1), 1e-3 * gscale):
            continue  # piecewise-smooth loss: a clamp switch or cell crossing inside +-eps
        kept += 1
        rels.append(abs(g - f1) / max(abs(f1), 1e-2 * gscale))
    print("  %d of 32 entries smooth; rel err median %.2e, max %.2e" % (kept, np.median(rels), max(rels)))
    assert kept >= 20
    assert max(rels) <= 1e-4


def test_torch_autograd_function(hip):
    """The LibTorch autograd function gives the same gradients as the C ABI calls."""
    torch = hip
    c, u, rho, dx, dt = small_scene(6, cfl=0.6)
    _case("torch autograd function, ragged scene")
    ext = diff.torch_extension()
    assert ext is not None, "hns_torch extension not built"
    t = _topo(c, dx)
    target = torch.from_numpy((np.roll(rho, 11) * 0.9).astype(np.float32)).cuda()
    v0 = _soa(torch, u).requires_grad_(True)
    r0 = torch.from_numpy(rho).cuda().requires_grad_(True)
    v, r = diff.diff_steps(t, v0, r0, 3, dt, dx, 10, _omega(dx), 2)
    loss = ((r - target) ** 2).sum() + 0.1 * (v ** 2).sum()
    loss.backward()
    # the same through the C ABI
    ud, rd = v0.detach().clone(), r0.detach().clone()
    vo, ro = torch.empty_like(ud), torch.empty_like(rd)
    tape = diff.forward(t, ud, rd, 3, dt, dx, 10, _omega(dx), 2, vo, ro)
    gv0, gr0 = torch.empty_like(ud), torch.empty_like(rd)
    tape.backward((0.2 * vo).contiguous(), (2.0 * (ro - target)).contiguous(), gv0, gr0)
    tape.close()
    assert torch.allclose(v.detach(), vo) and torch.allclose(r.detach(), ro)
    ev = float((v0.grad - gv0).norm() / gv0.norm())
    er = float((r0.grad - gr0).norm() / gr0.norm())
    print("  autograd vs C ABI rel diff: d/dv %.2e d/drho %.2e" % (ev, er))
    assert ev <= 1e-5 and er <= 1e-5
