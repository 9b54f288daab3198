"""Reverse mode of the differentiable step on the GPU (hnsk_diff_* in include/hns.h).

Step: u1 = advect_vector(u), div = divergence(u1), p = K red-black Gauss-Seidel iterations from 0,
u2 = u1 - grad p, rho' = advect_scalars(rho; u2), u' = u2 (src/Cuda/Kernel.cu:349-426, :432-453,
:459-489, :495-541, :163-289). Gradients are exact transposes of those kernels, computed on the
device with checkpoints every k steps. `diff_steps` exposes the same as a torch autograd function:
the C++ torch::autograd::Function of the in-tree hns_torch extension (LibTorch-ROCm host code over
the same C ABI).
"""
from __future__ import annotations

import ctypes as C
import importlib
import os
import sys

from . import _check, lib


def _ptr(t):
    if t is None:
        return None
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("expected a contiguous device tensor")
    return C.c_void_p(t.data_ptr())


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Tape:
    """hns_diff_tape: forward states at checkpoints."""

    def __init__(self, h):
        self._h = h

    def info(self):
        n, b, m = C.c_int32(), C.c_int64(), C.c_int64()
        _check(lib().hns_diff_tape_info(self._h, C.byref(n), C.byref(b), C.byref(m)))
        return {"checkpoints": n.value, "bytes": b.value, "recompute_mismatches": m.value}

    def backward(self, gvel_out, grho_out, gvel0, grho0):
        _check(lib().hnsk_diff_backward(self._h, _ptr(gvel_out), _ptr(grho_out), _ptr(gvel0), _ptr(grho0), _stream()))

    def close(self):
        if self._h:
            _check(lib().hns_diff_tape_destroy(C.byref(self._h)))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def forward(topo, vel0, rho0, steps, dt, dx, iterations, omega, checkpoint_every, vel_out=None, rho_out=None) -> Tape:
    h = C.c_void_p()
    _check(lib().hnsk_diff_forward(topo.handle, _ptr(vel0), _ptr(rho0), int(steps), float(dt), float(dx),
                                   int(iterations), float(omega), int(checkpoint_every), _ptr(vel_out), _ptr(rho_out),
                                   C.byref(h), _stream()))
    return Tape(h)


def advect_scalars_adj(topo, vel, field, gout, gin, gvel, dt, inv_dx):
    _check(lib().hnsk_advect_scalars_adj(topo.handle, _ptr(vel), _ptr(field), _ptr(gout), _ptr(gin), _ptr(gvel),
                                         float(dt), float(inv_dx), _stream()))


def advect_vector_adj(topo, vel, gout, gvel, dt, inv_dx):
    _check(lib().hnsk_advect_vector_adj(topo.handle, _ptr(vel), _ptr(gout), _ptr(gvel), float(dt), float(inv_dx),
                                        _stream()))


def rbgs_adj(topo, g_cs, gdiv_cs, dx, iterations, omega):
    _check(lib().hnsk_rbgs_adj(topo.handle, _ptr(g_cs), _ptr(gdiv_cs), float(dx), int(iterations), float(omega),
                               _stream()))


_EXT = None


def torch_extension():
    """The hns_torch C++ extension (torch::autograd::Function over libhns), or None if not built."""
    global _EXT
    if _EXT is None:
        here = os.path.dirname(os.path.abspath(__file__))
        if here not in sys.path:
            sys.path.insert(0, here)
        try:
            _EXT = importlib.import_module("hns_torch")
        except ImportError:
            _EXT = False
    return _EXT or None


def diff_steps(topo, vel0, rho0, steps, dt, dx, iterations, omega, checkpoint_every=4):
    """(u_S, rho_S) after `steps` differentiable steps, differentiable w.r.t. vel0 (SoA (3, N)) and
    rho0 (N,) through torch.autograd (raises ImportError when hns_torch is not built)."""
    import torch
    ext = torch_extension()
    if ext is None:
        raise ImportError("hns_torch is not built; run `python -m hnanosolver_amd.build --torch`")
    lib()  # libhns is loaded first so the extension's dependency resolves to the in-tree copy
    return ext.diff_steps(vel0, rho0, int(topo.handle.value), int(steps), float(dt), float(dx), int(iterations),
                          float(omega), int(checkpoint_every), int(torch.cuda.current_stream().cuda_stream))
