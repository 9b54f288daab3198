"""CPU pins of the FP64 host tape (the reverse mode of the differentiable step).

Linear pieces (divergence, gradient, K RBGS iterations, advect_scalars w.r.t. the field) pass the
dot-product test <A x, y> = <x, A^T y> to 1e-12 in FP64; the nonlinear advection adjoints match
FP64 central differences of the forward kernels; the loss gradient over several steps matches
FP64 central finite differences (eps 1e-3, the differentiable-control setting) at random voxels."""
import numpy as np
import pytest

import oracle

from _helpers import canonical_coords, ragged_origins


def small_scene(seed=0, cfl=0.8):
    c = canonical_coords(ragged_origins(seed)[:40])
    x = c.astype(np.float64)
    rng = np.random.default_rng(seed)
    u = np.stack([np.sin(0.2 * x[:, 1] + 0.3), np.cos(0.17 * x[:, 0]), 0.5 * np.sin(0.1 * x[:, 0] + 0.2 * x[:, 1])], -1)
    u += 0.2 * rng.standard_normal(u.shape)
    dx, dt = 0.1, 0.05
    u *= cfl * dx / dt / np.abs(u).max()
    rho = np.exp(-((x - x.mean(0)) ** 2).sum(1) / (2 * 6.0 ** 2)) + 0.1 * np.sin(0.3 * x[:, 2])
    return c, u, rho, dx, dt


def dot(a, b):
    return float(np.dot(np.ravel(a), np.ravel(b)))


def test_dot_product_divergence_and_gradient():
    c, u, rho, dx, dt = small_scene(0)
    o = oracle.Oracle(c, np.float64)
    rng = np.random.default_rng(1)
    x = rng.standard_normal((len(c), 3))
    y = rng.standard_normal(len(c))
    inv_dx = 1.0 / np.float32(dx)
    lhs = dot(o.divergence(x, inv_dx), y)
    rhs = dot(x, o.divergence_adj(y, inv_dx))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    p = rng.standard_normal(len(c))
    g = -o.subtract_gradient(np.zeros((len(c), 3)), p, inv_dx)   # = grad p
    lhs = dot(g, x)
    rhs = -dot(p, o.gradient_adj(x, inv_dx))  # gradient_adj is the adjoint of u1 - grad p
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


@pytest.mark.parametrize("K", [1, 2, 7])
def test_dot_product_rbgs(K):
    c, u, rho, dx, dt = small_scene(1)
    o = oracle.Oracle(c, np.float64)
    rng = np.random.default_rng(2)
    div = rng.standard_normal(len(c))
    y = rng.standard_normal(len(c))
    omega = 2.0 / (1.0 + np.sin(np.float32(3.14159) * np.float32(dx)))
    p = np.zeros(len(c))
    for _ in range(K):
        p = o.rbgs(div, p, dx, 0, omega)
        p = o.rbgs(div, p, dx, 1, omega)
    lhs = dot(p, y)
    rhs = dot(div, o.rbgs_adj(y, dx, K, omega))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_dot_product_advect_scalars_field():
    c, u, rho, dx, dt = small_scene(2)
    o = oracle.Oracle(c, np.float64)
    rng = np.random.default_rng(3)
    f = rng.random(len(c))
    y = rng.standard_normal(len(c))
    inv_dx = 1.0 / np.float32(dx)
    # piecewise linear in f: the clamp branch is fixed for a small perturbation, so check the
    # directional derivative (exact for a linear piece) against the adjoint
    v = rng.standard_normal(len(c))
    eps = 1e-7
    jv = (o.advect_scalars(u, [f + eps * v], dt, inv_dx)[0] - o.advect_scalars(u, [f - eps * v], dt, inv_dx)[0]) / (2 * eps)
    gin, _ = o.advect_scalars_adj(u, f, y, dt, inv_dx)
    assert abs(dot(jv, y) - dot(v, gin)) <= 1e-6 * abs(dot(v, gin))


def test_advection_velocity_adjoints_match_central_differences():
    c, u, rho, dx, dt = small_scene(3)
    o = oracle.Oracle(c, np.float64)
    rng = np.random.default_rng(4)
    inv_dx = 1.0 / np.float32(dx)
    y = rng.standard_normal(len(c))
    Y = rng.standard_normal((len(c), 3))
    v = rng.standard_normal((len(c), 3))
    eps = 1e-6
    # advect_scalars w.r.t. the velocity
    jv = (o.advect_scalars(u + eps * v, [rho], dt, inv_dx)[0] - o.advect_scalars(u - eps * v, [rho], dt, inv_dx)[0]) / (2 * eps)
    _, gvel = o.advect_scalars_adj(u, rho, y, dt, inv_dx)
    assert abs(dot(jv, y) - dot(v, gvel)) <= 1e-5 * abs(dot(v, gvel))
    # advect_vector (velocity transports itself)
    jv = (o.advect_vector(u + eps * v, dt, inv_dx) - o.advect_vector(u - eps * v, dt, inv_dx)) / (2 * eps)
    g = o.advect_vector_adj(u, Y, dt, inv_dx)
    assert abs(dot(jv, Y) - dot(v, g)) <= 1e-5 * abs(dot(v, g))


def test_loss_gradient_matches_fp64_central_differences():
    """dL/du0 of L = ||rho_S - target||^2 after 4 steps (advect, 20 RBGS, project, advect density)
    against FP64 central differences with eps = 1e-3 at random velocity components."""
    c, u, rho, dx, dt = small_scene(5, cfl=0.6)
    o = oracle.Oracle(c, np.float64)
    rng = np.random.default_rng(6)
    target = np.roll(rho, 37) * 0.8
    S, K = 4, 20
    omega = float(np.float32(2.0 / (1.0 + np.sin(np.float32(3.14159) * np.float32(dx)))))
    loss, gv, gr = o.diff_loss_grad(u, rho, target, S, dt, dx, K, omega)

    def L(uu, rr):
        _, r = o.diff_forward(uu, rr, S, dt, dx, K, omega)
        return float(((r - target) ** 2).sum())

    assert L(u, rho) == pytest.approx(loss, rel=1e-12)
    # sample voxels where the gradient is not negligible
    cand = np.argsort(-np.abs(gv).max(1))[:400]
    picks = rng.choice(cand, 24, replace=False)
    eps = 1e-3

    def fd_at(i, k, e):
        up, um = u.copy(), u.copy()
        up[i, k] += e
        um[i, k] -= e
        return (L(up, rho) - L(um, rho)) / (2 * e)

    kept, rels = 0, []
    for i in picks:
        k = int(np.argmax(np.abs(gv[i])))
        f1, f2 = fd_at(i, k, eps), fd_at(i, k, eps / 2)
        # a clamp switch or a trilinear cell crossing inside +-eps makes L only piecewise smooth:
        # there the central difference averages two slopes and halving eps moves it
        if abs(f1 - f2) > 1e-5 * abs(f1):
            continue
        kept += 1
        rels.append(abs(f1 - gv[i, k]) / abs(f1))
    print("FD vs tape at %d smooth samples of 24: median rel %.2e, max rel %.2e" % (kept, np.median(rels), max(rels)))
    assert kept >= 16
    assert max(rels) <= 1e-4
    # the density gradient, the same way
    def fdr(j, e):
        rp, rm = rho.copy(), rho.copy()
        rp[j] += e
        rm[j] -= e
        return (L(u, rp) - L(u, rm)) / (2 * e)

    kept = 0
    for j in rng.choice(np.argsort(-np.abs(gr))[:400], 12, replace=False):
        f1, f2 = fdr(j, eps), fdr(j, eps / 2)
        if abs(f1 - f2) > 1e-5 * abs(f1):
            continue
        kept += 1
        assert abs(f1 - gr[j]) <= 1e-4 * abs(f1)
    assert kept >= 6
