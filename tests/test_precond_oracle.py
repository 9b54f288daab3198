"""CPU checks of the dense restatement of the learned preconditioner (oracle/precond_net.py):
brute-force sparse convolution on scattered leaves, adjoint (dot-product) identities, central
finite differences of the loss in the weights, and the closed forms of the operator."""
import numpy as np
import pytest

from oracle import precond_net as PN
from hnanosolver_amd import precond as P


def _origins(seed, n=9):
    rng = np.random.default_rng(seed)
    cand = [(x, y, z) for x in range(0, 32, 8) for y in range(0, 24, 8) for z in range(0, 24, 8)]
    idx = rng.choice(len(cand), n, replace=False)
    return np.array([cand[i] for i in sorted(idx)], np.int64)


def test_layer_table_matches_product():
    assert PN.LAYERS == P.LAYERS and PN.N_PARAMS == P.N_PARAMS == 217280


def test_dense_conv_equals_brute_force_sparse_conv():
    """conv over the dense box (zeros outside active cells) == explicit per-cell sum over the 27
    taps of a dictionary of active cells, at every level."""
    org = _origins(1)
    net = PN.NetOracle(org)
    rng = np.random.default_rng(2)
    for lev, (ci, co) in ((0, (3, 4)), (1, (5, 2)), (2, (4, 3))):
        S = 8 >> lev
        n = len(org) * S ** 3
        x = rng.standard_normal((n, ci))
        W = rng.standard_normal((27, ci, co))
        got = net.sparse(lev, net.conv(lev, net.dense(lev, x), W))
        cells = {tuple(c): i for i, c in enumerate(net.index[lev])}
        ref = np.zeros((n, co))
        for i, c in enumerate(net.index[lev]):
            for t, d in enumerate(PN.taps()):
                j = cells.get((c[0] + d[0], c[1] + d[1], c[2] + d[2]))
                if j is not None:
                    ref[i] += x[j] @ W[t]
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_adjoint_identities():
    org = _origins(3)
    net = PN.NetOracle(org)
    rng = np.random.default_rng(4)
    for lev in range(3):
        m = net.masks[lev][..., None]
        x = rng.standard_normal(m.shape[:3] + (3,)) * m
        g = rng.standard_normal(m.shape[:3] + (2,)) * m
        W = rng.standard_normal((27, 3, 2))
        lhs = np.sum(net.conv(lev, x, W) * g)
        assert abs(lhs - np.sum(x * net.conv_adj_x(lev, g, W))) < 1e-9 * abs(lhs)
        assert abs(lhs - np.sum(W * net.conv_adj_w(lev, x, g))) < 1e-9 * abs(lhs)
    x = rng.standard_normal(net.masks[0].shape + (2,))
    y = rng.standard_normal(net.masks[1].shape + (2,))
    assert abs(np.sum(net.pool(x) * y) - np.sum(x * net.up(y) / 8)) < 1e-10


def test_operator_closed_form():
    org = _origins(5)
    net = PN.NetOracle(org)
    n = len(org) * 512
    for i in (0, 7, 63, 300, n - 1):
        e = np.zeros(n)
        e[i] = 1.0
        Ae = net.apply_A(e)
        c = net.index[0][i]
        assert Ae[i] == 6.0
        nb = {tuple(net.index[0][j]) for j in np.nonzero(Ae == -1.0)[0]}
        want = {(c[0] + d[0], c[1] + d[1], c[2] + d[2]) for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0),
                                                                  (0, 0, 1), (0, 0, -1))}
        assert nb == want & {tuple(v) for v in net.index[0]}
        assert np.count_nonzero(Ae) == 1 + len(nb)


def test_positive_homogeneity_and_bf16_rounding():
    org = _origins(6, 5)
    net = PN.NetOracle(org)
    w = P.init_params(3)
    r = np.random.default_rng(7).standard_normal(len(org) * 512)
    z1, _ = net.forward(w, r)
    z2, _ = net.forward(w, 2.5 * r)
    np.testing.assert_allclose(z2, 2.5 * z1, rtol=1e-12, atol=1e-12 * np.abs(z1).max())
    # brute force: the nearest of all floats whose low 16 bits are zero, ties to the even one
    a = np.array([1.0, 1.00390625, 1.005859375, -3.0e-5, 123456.0, 3.0e38, 1.01171875], np.float32)
    for v, got in zip(a, PN.bf16(a)):
        u = int(np.array([v]).view(np.uint32)[0]) & 0xFFFF0000
        lo, hi = np.array([u, u + 0x10000], np.uint32).view(np.float32).astype(np.float64)
        dl, dh = abs(float(v) - lo), abs(float(v) - hi)
        want = lo if (dl < dh or (dl == dh and (u >> 16) % 2 == 0)) else hi
        assert float(got) == want
    assert list(PN.bf16(a[:3])) == [1.0, 1.0, 1.0078125]


def test_loss_gradient_matches_central_differences():
    """FP64 backward of ||r - A z||^2/||r||^2 against central differences at 24 weights spread over
    all layers; samples whose FD moves across a relu kink (FD at eps and eps/2 disagree) are skipped."""
    org = _origins(8, 4)
    net = PN.NetOracle(org)
    w = P.init_params(11).astype(np.float64) * 0.7
    r = np.random.default_rng(12).standard_normal(len(org) * 512)
    L0, g, _ = net.loss(w, r)
    rng = np.random.default_rng(13)
    checked, worst = 0, 0.0
    for name, (o, (t, ci, co)) in PN.OFFSETS.items():
        for i in rng.choice(t * ci * co, 4 if name != "out" else 3, replace=False):
            k = o + i

            def fd(eps):
                wp, wm = w.copy(), w.copy()
                wp[k] += eps
                wm[k] -= eps
                return (net.loss(wp, r)[0] - net.loss(wm, r)[0]) / (2 * eps)
            f1, f2 = fd(1e-6), fd(5e-7)
            scale = max(abs(g[k]), 1e-3 * np.abs(g).max())
            if abs(f1 - f2) > 1e-4 * scale:
                continue
            checked += 1
            worst = max(worst, abs(f1 - g[k]) / scale)
    assert checked >= 20, checked
    assert worst < 1e-4, worst
