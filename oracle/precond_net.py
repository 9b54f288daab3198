"""TEST INFRASTRUCTURE — dense numpy restatement of the learned pressure preconditioner.

Only tests/ may import this module; the product path never does. The network has no reference
counterpart (the reference solves pressure with a fixed red-black Gauss-Seidel count,
src/Cuda/HNanoSolver.cu:364-373). It is restated here on dense arrays: a sparse 3x3x3 convolution
over active leaves equals a dense convolution over the bounding box with zeros outside the active
cells, masked back to the active cells. That makes it checkable against plain dense arithmetic
(FP64), finite differences on the weights, and the sparse GPU kernels.

Network (all convolutions without bias, so z(a r) = a z(r) for a > 0):
    level 0: voxels, level 1: 2^3 voxels per cell, level 2: 4^3 voxels per cell (4^3/2^3 cells per leaf)
    h0 = relu(conv3(r; c0))                     1 -> 16, level 0
    h1 = relu(conv3(pool(h0); c1))             16 -> 32, level 1
    h2 = relu(conv3(pool(h1); c2))             32 -> 64, level 2
    h3 = relu(conv3(h2; c3))                   64 -> 64, level 2
    s1 = h1 + up(conv1(h3; u2))                64 -> 32, level 2 -> 1
    h5 = relu(conv3(s1; c5))                   32 -> 32, level 1
    s0 = h0 + up(conv1(h5; u1))                32 -> 16, level 1 -> 0
    h6 = relu(conv3(s0; c6))                   16 -> 16, level 0
    z  = h6 . w_out                            16 -> 1
conv3(x; W)[c, o] = sum_{t, i} W[t, i, o] x[c + d(t), i], d(t) = (t//9 - 1, (t//3)%3 - 1, t%3 - 1);
pool = mean of the 8 children; up = copy of the parent. With bf16=True every stored activation
and the weights are rounded to bfloat16 (round to nearest even) where the GPU kernels store them.
"""
from __future__ import annotations

import numpy as np

# (name, taps, cin, cout, level)
LAYERS = [("c0", 27, 1, 16, 0), ("c1", 27, 16, 32, 1), ("c2", 27, 32, 64, 2), ("c3", 27, 64, 64, 2),
          ("u2", 1, 64, 32, 2), ("c5", 27, 32, 32, 1), ("u1", 1, 32, 16, 1), ("c6", 27, 16, 16, 0),
          ("out", 1, 16, 1, 0)]


def layer_offsets():
    off, o = {}, 0
    for name, t, ci, co, _ in LAYERS:
        off[name] = (o, (t, ci, co))
        o += t * ci * co
    return off, o


OFFSETS, N_PARAMS = layer_offsets()


def unpack(params):
    return {k: params[o:o + t * ci * co].reshape(t, ci, co) for k, (o, (t, ci, co)) in OFFSETS.items()}


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def taps():
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


class NetOracle:
    """Dense restatement over the leaves `origins` (int [L, 3], multiples of 8, in the sparse
    cell order: cell = leaf * S^3 + (cx * S + cy) * S + cz with S = 8 >> level)."""

    def __init__(self, origins, dtype=np.float64, bf16_round=False):
        self.origins = np.asarray(origins, np.int64)
        self.dtype = dtype
        self.round = bf16_round
        self.lo = self.origins.min(0)
        self.dims = (self.origins.max(0) - self.lo) // 8 + 1  # in leaves
        self.L = len(self.origins)
        self.masks, self.index = [], []
        for lev in range(3):
            S = 8 >> lev
            D = tuple(self.dims * S)
            m = np.zeros(D, bool)
            idx = np.zeros((self.L * S ** 3, 3), np.int64)
            g = np.arange(S)
            cx, cy, cz = np.meshgrid(g, g, g, indexing="ij")
            local = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], -1)
            for l, o in enumerate(self.origins):
                base = (o - self.lo) // 8 * S
                idx[l * S ** 3:(l + 1) * S ** 3] = base + local
            m[idx[:, 0], idx[:, 1], idx[:, 2]] = True
            self.masks.append(m)
            self.index.append(idx)

    # ---- sparse <-> dense
    def dense(self, lev, x):
        x = np.asarray(x, self.dtype)
        x2 = x.reshape(len(self.index[lev]), -1)
        out = np.zeros(self.masks[lev].shape + (x2.shape[1],), self.dtype)
        i = self.index[lev]
        out[i[:, 0], i[:, 1], i[:, 2]] = x2
        return out

    def sparse(self, lev, d):
        i = self.index[lev]
        return d[i[:, 0], i[:, 1], i[:, 2]]

    def _r(self, a):
        return bf16(a).astype(self.dtype) if self.round else a

    # ---- dense ops (inactive cells are zero on input and output)
    def conv(self, lev, x, W):
        """x dense [X,Y,Z,ci]; W [taps, ci, co]."""
        m = self.masks[lev]
        if W.shape[0] == 1:
            return (x @ W[0]) * m[..., None]
        xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
        X, Y, Z = m.shape
        out = np.zeros((X, Y, Z, W.shape[2]), self.dtype)
        for t, (dx, dy, dz) in enumerate(taps()):
            out += xp[1 + dx:1 + dx + X, 1 + dy:1 + dy + Y, 1 + dz:1 + dz + Z] @ W[t]
        return out * m[..., None]

    def conv_adj_x(self, lev, g, W):
        """d/dx of sum(g * conv(x; W)) (inactive cells of g are zero)."""
        m = self.masks[lev]
        g = g * m[..., None]
        if W.shape[0] == 1:
            return (g @ W[0].T) * m[..., None]
        gp = np.pad(g, ((1, 1), (1, 1), (1, 1), (0, 0)))
        X, Y, Z = m.shape
        out = np.zeros((X, Y, Z, W.shape[1]), self.dtype)
        for t, (dx, dy, dz) in enumerate(taps()):
            out += gp[1 - dx:1 - dx + X, 1 - dy:1 - dy + Y, 1 - dz:1 - dz + Z] @ W[t].T
        return out * m[..., None]

    def conv_adj_w(self, lev, x, g):
        """d/dW of sum(g * conv(x; W)), [taps, ci, co]."""
        m = self.masks[lev]
        g = (g * m[..., None]).reshape(-1, g.shape[-1])
        X, Y, Z = m.shape
        xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
        out = []
        for dx, dy, dz in taps():
            xs = xp[1 + dx:1 + dx + X, 1 + dy:1 + dy + Y, 1 + dz:1 + dz + Z].reshape(-1, x.shape[-1])
            out.append(xs.T @ g)
        return np.stack(out)

    def conv1_adj_w(self, lev, x, g):
        m = self.masks[lev]
        return ((x.reshape(-1, x.shape[-1])).T @ (g * m[..., None]).reshape(-1, g.shape[-1]))[None]

    @staticmethod
    def pool(x):
        X, Y, Z, C = x.shape
        return x.reshape(X // 2, 2, Y // 2, 2, Z // 2, 2, C).mean((1, 3, 5))

    @staticmethod
    def up(x):
        return x.repeat(2, 0).repeat(2, 1).repeat(2, 2)

    # ---- network
    def forward(self, params, r):
        """r: level-0 sparse [n0] (leaf-dense order). Returns (z sparse [n0], cache)."""
        if self.round:
            P = {k: bf16(np.asarray(v, np.float32)).astype(self.dtype) for k, v in unpack(params).items()}
        else:
            P = {k: np.asarray(v, self.dtype) for k, v in unpack(params).items()}
        relu = lambda a: np.maximum(a, 0)  # noqa: E731
        R = self._r
        x0 = R(self.dense(0, r))
        h0 = R(relu(self.conv(0, x0, P["c0"])))
        p1 = R(self.pool(h0))
        h1 = R(relu(self.conv(1, p1, P["c1"])))
        p2 = R(self.pool(h1))
        h2 = R(relu(self.conv(2, p2, P["c2"])))
        h3 = R(relu(self.conv(2, h2, P["c3"])))
        g2 = R(self.conv(2, h3, P["u2"]))
        s1 = R(h1 + self.up(g2) * self.masks[1][..., None])
        h5 = R(relu(self.conv(1, s1, P["c5"])))
        g1 = R(self.conv(1, h5, P["u1"]))
        s0 = R(h0 + self.up(g1) * self.masks[0][..., None])
        h6 = relu(self.conv(0, s0, P["c6"]))
        z = self.conv(0, R(h6), P["out"])
        cache = dict(P=P, x0=x0, h0=h0, p1=p1, h1=h1, p2=p2, h2=h2, h3=h3, s1=s1, h5=h5, s0=s0, h6=R(h6))
        return self.sparse(0, z)[:, 0], cache

    def backward(self, cache, dz):
        """d/dparams of sum(dz * z) (FP64 arithmetic on the cached forward values)."""
        c, P = cache, cache["P"]
        m0, m1 = self.masks[0][..., None], self.masks[1][..., None]
        g = {}
        dzd = self.dense(0, dz)
        g["out"] = self.conv1_adj_w(0, c["h6"], dzd)
        dh6 = self.conv_adj_x(0, dzd, P["out"]) * (c["h6"] > 0)
        g["c6"] = self.conv_adj_w(0, c["s0"], dh6)
        ds0 = self.conv_adj_x(0, dh6, P["c6"])
        dh0 = ds0.copy()
        dg1 = self.pool(ds0 * m0) * 8
        g["u1"] = self.conv1_adj_w(1, c["h5"], dg1)
        dh5 = self.conv_adj_x(1, dg1, P["u1"]) * (c["h5"] > 0)
        g["c5"] = self.conv_adj_w(1, c["s1"], dh5)
        ds1 = self.conv_adj_x(1, dh5, P["c5"])
        dh1 = ds1.copy()
        dg2 = self.pool(ds1 * m1) * 8
        g["u2"] = self.conv1_adj_w(2, c["h3"], dg2)
        dh3 = self.conv_adj_x(2, dg2, P["u2"]) * (c["h3"] > 0)
        g["c3"] = self.conv_adj_w(2, c["h2"], dh3)
        dh2 = self.conv_adj_x(2, dh3, P["c3"]) * (c["h2"] > 0)
        g["c2"] = self.conv_adj_w(2, c["p2"], dh2)
        dp2 = self.conv_adj_x(2, dh2, P["c2"])
        dh1 = (dh1 + self.up(dp2) / 8 * m1) * (c["h1"] > 0)
        g["c1"] = self.conv_adj_w(1, c["p1"], dh1)
        dp1 = self.conv_adj_x(1, dh1, P["c1"])
        dh0 = (dh0 + self.up(dp1) / 8 * m0) * (c["h0"] > 0)
        g["c0"] = self.conv_adj_w(0, c["x0"], dh0)
        out = np.zeros(N_PARAMS, np.float64)
        for k, (o, (t, ci, co)) in OFFSETS.items():
            out[o:o + t * ci * co] = g[k].reshape(-1)
        return out

    # ---- Poisson operator (A = 6 I - face neighbours, zero outside the active voxels)
    def apply_A(self, x):
        d = self.dense(0, x)[..., 0]
        dp = np.pad(d, 1)
        X, Y, Z = d.shape
        s = (dp[2:, 1:-1, 1:-1] + dp[:-2, 1:-1, 1:-1] + dp[1:-1, 2:, 1:-1] + dp[1:-1, :-2, 1:-1] +
             dp[1:-1, 1:-1, 2:] + dp[1:-1, 1:-1, :-2])
        return self.sparse(0, (6 * d - s)[..., None])[:, 0]

    def loss(self, params, r):
        """||r - A z||^2 / ||r||^2 and its gradient in the parameters."""
        z, cache = self.forward(params, r)
        u = r - self.apply_A(z)
        rr = float(np.dot(r, r))
        L = float(np.dot(u, u)) / rr
        dz = -2.0 * self.apply_A(u) / rr
        return L, self.backward(cache, dz), z
