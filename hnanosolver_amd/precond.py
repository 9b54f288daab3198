"""Learned pressure preconditioner: weights, GPU training, and the preconditioned solve.

The network (a sparse 3-level 3x3x3 convolutional net over the active leaves, 16/32/64 channels,
BF16 weights and activations with FP32 accumulation on MFMA) maps a residual r of the pressure
system A p = b (A = 6 I - face neighbours, b = -div dx^2) to a correction z ~ A^-1 r. It serves as
the preconditioner of a flexible PCG (HNS_SOLVER_PCG_LEARNED); plain red-black Gauss-Seidel remains
the fallback and the judge of the residual. There is no reference counterpart: the reference
projects with a fixed RBGS count (src/Cuda/HNanoSolver.cu:364-373).

Training runs on the GPU with the library's own forward and backward kernels. Pairs (r_k, e_k)
come from the plain RBGS solver: r_k = b - A p_k after k iterations, e_k = p* - p_k with p* the
converged solution, so A e_k = r_k. The loss is the error of the pair in the energy norm,
    ||z(r) - e||_A^2 / ||e||_A^2 = (z.A z - 2 z.r) / (r.e) + 1,
the quantity that bounds the convergence of (flexible) PCG; at run time it needs r and the scalar
r.e of the pair only.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional, Sequence

import numpy as np

from . import SOLVERS, _check, hns_solve_options, hns_solve_stats, lib

# (name, taps, cin, cout, level) in parameter order; W[tap][cin][cout] per layer
LAYERS = [("c0", 27, 1, 16, 0), ("c1", 27, 16, 32, 1), ("c2", 27, 32, 64, 2), ("c3", 27, 64, 64, 2),
          ("u2", 1, 64, 32, 2), ("c5", 27, 32, 32, 1), ("u1", 1, 32, 16, 1), ("c6", 27, 16, 16, 0),
          ("out", 1, 16, 1, 0)]
N_PARAMS = sum(t * ci * co for _, t, ci, co, _ in LAYERS)
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "precond_weights.npz")
POISSON_WORK = 6148  # HNS_POISSON_WORK_DOUBLES


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights (fan-in = taps * cin), float32."""
    rng = np.random.default_rng(seed)
    out = np.empty(N_PARAMS, np.float32)
    o = 0
    for name, t, ci, co, _ in LAYERS:
        std = np.sqrt(2.0 / (t * ci)) if name != "out" else np.sqrt(1.0 / ci)
        out[o:o + t * ci * co] = rng.standard_normal(t * ci * co) * std
        o += t * ci * co
    return out


def load_weights(path: str = WEIGHTS_PATH) -> np.ndarray:
    with np.load(path, allow_pickle=False) as f:
        w = np.ascontiguousarray(f["params"], np.float32)
    if w.shape != (N_PARAMS,):
        raise ValueError("weight file holds %s parameters, expected %d" % (w.shape, N_PARAMS))
    return w


def save_weights(params, path: str = WEIGHTS_PATH, **meta):
    np.savez(path, params=np.asarray(params, np.float32), **{k: np.asarray(v) for k, v in meta.items()})


def _ptr(t):
    if t is None:
        return None
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("expected a contiguous device tensor")
    return C.c_void_p(t.data_ptr())


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Preconditioner:
    """hns_precond handle. `params` is a host array (copied) or a float32 device tensor (shared;
    call refresh() after changing it)."""

    def __init__(self, params=None):
        if lib().hns_precond_num_params() != N_PARAMS:
            raise RuntimeError("libhns and this module disagree on the network size")
        h = C.c_void_p()
        if params is None:
            params = load_weights()
        if isinstance(params, np.ndarray):
            w = np.ascontiguousarray(params, np.float32)
            _check(lib().hns_precond_create(C.c_void_p(w.ctypes.data), N_PARAMS, C.byref(h)))
            self.device_params = None
        else:
            _check(lib().hns_precond_create_device(_ptr(params), N_PARAMS, C.byref(h)))
            self.device_params = params
        self._h = h

    @property
    def handle(self):
        if not self._h:
            raise RuntimeError("preconditioner destroyed")
        return self._h

    def refresh(self):
        _check(lib().hns_precond_refresh(self.handle, _stream()))

    def apply(self, topo, r_cs, z_cs, keep: bool = False):
        _check(lib().hnsk_precond_apply(topo.handle, self.handle, _ptr(r_cs), _ptr(z_cs), int(keep), _stream()))

    def backward(self, topo, r_cs, dz_cs, grad):
        _check(lib().hnsk_precond_backward(topo.handle, self.handle, _ptr(r_cs), _ptr(dz_cs), _ptr(grad), _stream()))

    def solve(self, topo, div_cs, p_cs, dx, tol=1e-5, max_iters=10000):
        """Flexible PCG from p = 0 to ||div - L p|| / ||div|| <= tol; returns hns_solve_stats."""
        o = hns_solve_options(C.sizeof(hns_solve_options), SOLVERS["pcg_learned"], float(tol), int(max_iters), 0)
        s = hns_solve_stats(C.sizeof(hns_solve_stats), 0, -1.0, 0.0)
        _check(lib().hnsk_pressure_solve_learned(topo.handle, self.handle, _ptr(div_cs), _ptr(p_cs), dx, C.byref(o),
                                                 C.byref(s), _stream()))
        return s

    def close(self):
        if self._h:
            _check(lib().hns_precond_destroy(C.byref(self._h)))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_apply(topo, x_cs, y_cs, a, b, out_cs, work):
    """out = a*y + b*(A x); work (float64 device, POISSON_WORK) gets ||out||^2, ||y||^2, out.x in
    work[0], work[1], work[2]."""
    _check(lib().hnsk_poisson_apply(topo.handle, _ptr(x_cs), _ptr(y_cs), float(a), float(b), _ptr(out_cs),
                                    _ptr(work), _stream()))


class Trainer:
    """Adam on the device-resident FP32 weights, gradients from the library's backward kernels."""

    def __init__(self, params: np.ndarray, lr: float = 1e-3, betas=(0.9, 0.999), eps: float = 1e-8):
        import torch
        self.torch = torch
        self.params = torch.from_numpy(np.ascontiguousarray(params, np.float32)).cuda()
        self.grad = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.work = torch.zeros(POISSON_WORK + 3, dtype=torch.float64, device="cuda")
        self.pc = Preconditioner(self.params)
        self.lr, self.betas, self.eps, self.t = lr, betas, eps, 0
        self._bufs = {}

    def _buf(self, n):
        b = self._bufs.get(n)
        if b is None:
            t = self.torch
            b = self._bufs[n] = (t.empty(n, device="cuda"), t.empty(n, device="cuda"), t.empty(n, device="cuda"))
        return b

    def loss_grad(self, topo, r_cs, re: float, accumulate: bool = True) -> float:
        """Energy-norm loss (z.Az - 2 z.r)/(r.e) + 1 of one pair (r, e) given r.e; adds its gradient
        2 (A z - r)/(r.e) to self.grad."""
        z, u, dz = self._buf(r_cs.numel())
        self.pc.apply(topo, r_cs, z, keep=accumulate)
        poisson_apply(topo, z, r_cs, -2.0, 1.0, u, self.work)  # u = A z - 2 r; work[2] = z.u
        if accumulate:
            poisson_apply(topo, z, r_cs, -2.0 / re, 2.0 / re, dz, self.work[3:])  # dL/dz
            self.pc.backward(topo, r_cs, dz, self.grad)
        return float(self.work[2].item()) / re + 1.0

    def step(self):
        self.t += 1
        _check(lib().hnsk_precond_adam(N_PARAMS, _ptr(self.params), _ptr(self.grad), _ptr(self.m), _ptr(self.v),
                                       float(self.lr), float(self.betas[0]), float(self.betas[1]), float(self.eps),
                                       self.t, _stream()))
        self.grad.zero_()
        self.pc.refresh()

    def host_params(self) -> np.ndarray:
        return self.params.cpu().numpy().copy()


class Sample:
    """One topology with its training pairs: residuals r (colour-split float32 device tensors) and
    the energy norms r.e = ||e||_A^2 of their errors."""

    def __init__(self, topo, residuals: List, energies: List[float], dx: float, div_cs=None):
        self.topo, self.residuals, self.energies, self.dx, self.div_cs = topo, residuals, energies, dx, div_cs


def rbgs_pairs(topo, b_cs, dx, ks: Sequence[int], work, tol: float = 3e-6):
    """Pairs of the plain RBGS solver on A p = b (b = -rhs dx^2 for rhs = b_cs): for k in ks,
    r_k = b - A p_k and r_k.e_k with e_k = p* - p_k, p* the CG solution to `tol` or as close as
    FP32 gets in 4000 iterations (p* only normalises the loss; the pairs' residuals are far above
    tol)."""
    import torch

    from . import HNSError
    from . import kernels as K
    omega = float(np.float32(2.0 / (1.0 + np.sin(np.float32(3.14159) * np.float32(dx)))))
    pstar = torch.empty_like(b_cs)
    try:
        K.pressure_solve(topo, b_cs, pstar, dx, solver="cg", tol=tol, max_iters=4000)
    except HNSError:
        pass  # FP32 may stop short of tol; p* then normalises to that accuracy
    p = torch.zeros_like(b_cs)
    rs, es, done = [], [], 0
    for k in sorted(ks):
        while done < k:
            K.rbgs_sweep(topo, b_cs, p, dx, 0, omega, asynchronous=True)
            K.rbgs_sweep(topo, b_cs, p, dx, 1, omega, asynchronous=True)
            done += 1
        r = torch.empty_like(b_cs)
        poisson_apply(topo, p, b_cs, -dx * dx, -1.0, r, work)  # r = -rhs dx^2 - A p
        rs.append(r)
        es.append(_dot(topo, r, pstar, work) - _dot(topo, r, p, work))  # r.e, e = p* - p
    return rs, es


def _dot(topo, a, b, work) -> float:
    """a.b in FP64 through poisson_apply (out = a, x = b, out.x in work[2])."""
    import torch
    out = torch.empty_like(a)
    poisson_apply(topo, b, a, 1.0, 0.0, out, work)
    return float(work[2].item())


def rbgs_residual_pairs(topo, div_cs, dx, ks: Sequence[int], work) -> List:
    """Residuals r_k = b - A p_k (b = -div dx^2) of the plain RBGS iterates p_k, k in ks."""
    return rbgs_pairs(topo, div_cs, dx, ks, work)[0]


def scene_divergence(sc, topo):
    """Colour-split divergence of the scene's velocity after one advection and vorticity step
    (the right-hand side the step's projection sees)."""
    import torch

    from . import kernels as K
    u = torch.from_numpy(np.ascontiguousarray(sc.velocity.T)).cuda()
    a = torch.empty_like(u)
    K.advect_velocity(topo, u, a, sc.dt, 1.0 / sc.voxel_size, asynchronous=True)
    b = torch.empty_like(u)
    K.vorticity_forces(topo, a, b, sc.dt, 1.0 / sc.voxel_size, sc.params.get("vorticityScale", 0.0),
                       sc.params.get("factorScale", 1.0), asynchronous=True)
    d = torch.empty(sc.n, device="cuda")
    K.divergence(topo, b, d, 1.0 / sc.voxel_size, colour_split=True, asynchronous=True)
    return d


def make_samples(scene_fn, seeds: Sequence[int], ks=(0, 2, 8, 32, 128), noise: bool = True,
                 log=None) -> List[Sample]:
    """Training samples: for each seed, RBGS pairs of the scene's divergence and, with `noise`, of
    a seeded white-noise right-hand side on the same topology."""
    import torch

    from . import GridIndexedData, create_index_grid
    work = torch.zeros(POISSON_WORK, dtype=torch.float64, device="cuda")
    out = []
    import time
    for seed in seeds:
        t0 = time.perf_counter()
        sc = scene_fn(seed)
        if log:
            log("seed %d: scene built (%.2f s)" % (seed, time.perf_counter() - t0))
        topo = create_index_grid(GridIndexedData(sc.coords), sc.voxel_size)
        div = scene_divergence(sc, topo)
        rs, es = rbgs_pairs(topo, div, sc.voxel_size, ks, work)
        if log:
            log("seed %d: divergence pairs (%.2f s)" % (seed, time.perf_counter() - t0))
        if noise:
            g = torch.Generator(device="cuda")
            g.manual_seed(10007 * seed + 1)
            wn = torch.randn(sc.n, device="cuda", generator=g)
            r2, e2 = rbgs_pairs(topo, wn, sc.voxel_size, ks, work)
            rs, es = rs + r2, es + e2
        out.append(Sample(topo, rs, es, sc.voxel_size, div))
        if log:
            log("sample seed %d: N=%d, %d pairs" % (seed, sc.n, len(rs)))
    return out


def train(samples: List[Sample], steps: int, params: Optional[np.ndarray] = None, lr: float = 1e-3, seed: int = 0,
          log=None, log_every: int = 100):
    """Adam over random (sample, residual) draws; the learning rate decays to lr/10 (cosine).
    Returns (trained host params, list of (step, mean loss over the last log_every steps))."""
    rng = np.random.default_rng(seed)
    tr = Trainer(init_params(seed) if params is None else params, lr=lr)
    hist, acc, n = [], 0.0, 0
    for it in range(steps):
        tr.lr = lr * (0.55 + 0.45 * np.cos(np.pi * it / max(steps, 1)))
        s = samples[rng.integers(len(samples))]
        k = rng.integers(len(s.residuals))
        acc += tr.loss_grad(s.topo, s.residuals[k], s.energies[k])
        n += 1
        tr.step()
        if (it + 1) % log_every == 0 or it + 1 == steps:
            hist.append((it + 1, acc / n))
            if log:
                log("step %d loss %.4f" % (it + 1, acc / n))
            acc, n = 0.0, 0
    p = tr.host_params()
    tr.pc.close()
    return p, hist


def evaluate(params, scene_fn, seeds: Sequence[int], tol: float = 1e-5, log=None):
    """PCG (learned) against CG and plain RBGS to `tol` on held-out scenes: iterations and solve
    times per seed."""
    import torch

    from . import GridIndexedData, create_index_grid
    from . import kernels as K
    pc = Preconditioner(params)
    rows = []
    for seed in seeds:
        sc = scene_fn(seed)
        topo = create_index_grid(GridIndexedData(sc.coords), sc.voxel_size)
        div = scene_divergence(sc, topo)
        p = torch.empty_like(div)
        omega = float(np.float32(2.0 / (1.0 + np.sin(np.float32(3.14159) * np.float32(sc.voxel_size)))))
        row = {"seed": int(seed), "voxels": int(sc.n)}
        for name, fn in (("rbgs", lambda: K.pressure_solve(topo, div, p, sc.voxel_size, omega=omega, solver="rbgs_tol",
                                                          tol=tol, check_every=50)),
                         ("cg", lambda: K.pressure_solve(topo, div, p, sc.voxel_size, solver="cg", tol=tol)),
                         ("pcg_learned", lambda: pc.solve(topo, div, p, sc.voxel_size, tol=tol))):
            st = fn()
            row[name] = (int(st.iterations), float(st.solve_ms), float(st.rel_residual))
        rows.append(row)
        if log:
            log(str(row))
    pc.close()
    return rows


def _main(argv=None):
    import argparse

    import torch

    from . import scenes
    ap = argparse.ArgumentParser(description="train the learned pressure preconditioner on the GPU")
    ap.add_argument("--train-seeds", type=int, default=256)
    ap.add_argument("--eval-seeds", type=int, default=32)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=WEIGHTS_PATH)
    a = ap.parse_args(argv)
    torch.cuda.init()
    print("device ready", flush=True)
    samples = make_samples(scenes.precond128_2m, range(a.train_seeds), log=lambda m: print(m, flush=True))
    print("built %d samples" % len(samples), flush=True)
    p, hist = train(samples, a.steps, lr=a.lr, log=lambda m: print(m, flush=True), log_every=250)
    save_weights(p, a.out, train_seeds=a.train_seeds, steps=a.steps, loss=np.array(hist))
    print("saved", a.out, flush=True)
    del samples
    rows = evaluate(p, scenes.precond128_2m, range(10000, 10000 + a.eval_seeds), log=lambda m: print(m, flush=True))
    it = {k: np.mean([r[k][0] for r in rows]) for k in ("rbgs", "cg", "pcg_learned")}
    ms = {k: np.mean([r[k][1] for r in rows]) for k in ("rbgs", "cg", "pcg_learned")}
    print("held-out mean iterations %s, mean solve ms %s" % (it, ms), flush=True)


if __name__ == "__main__":
    _main()
