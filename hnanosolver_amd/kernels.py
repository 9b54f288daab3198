"""Device-array kernel API (hnsk_* in include/hns.h) over torch tensors on a HIP device.

Torch only supplies device memory and the current stream; every computation is a libhns
kernel. Velocity tensors are SoA (3, N) float32; scalars (N,) float32; pressure and divergence
for the RBGS sweep are colour-split (see to_colour_split).
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

from . import _check, lib

HNS_ERR_NOT_CONVERGED = -11  # a to-tolerance solve missed tol: iteration cap or breakdown (results written)


def _ptr(t):
    if t is None:
        return None
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("expected a contiguous device tensor")
    return C.c_void_p(t.data_ptr())


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def advect_velocity(topo, vel, out, dt, inv_dx, sdf=None, asynchronous=False):
    _check(lib().hnsk_advect_velocity(topo.handle, _ptr(vel), _ptr(out), dt, inv_dx, _ptr(sdf), _stream(),
                                      int(asynchronous)))


def advect_scalars(topo, vel, ins: Sequence, outs: Sequence, dt, inv_dx, sdf=None, asynchronous=False):
    m = len(ins)
    pin = (C.c_void_p * max(1, m))(*[t.data_ptr() for t in ins])
    pout = (C.c_void_p * max(1, m))(*[t.data_ptr() for t in outs])
    _check(lib().hnsk_advect_scalars(topo.handle, _ptr(vel), m, pin, pout, dt, inv_dx, _ptr(sdf), _stream(),
                                     int(asynchronous)))


def advect_scalar(topo, vel, field, out, dt, inv_dx, sdf=None, asynchronous=False):
    _check(lib().hnsk_advect_scalar(topo.handle, _ptr(vel), _ptr(field), _ptr(out), dt, inv_dx, _ptr(sdf),
                                    _stream(), int(asynchronous)))


def vorticity_forces(topo, vel, out, dt, inv_dx, scale, factor_scale, src_vel=None, gravity=0.0,
                     asynchronous=False):
    _check(lib().hnsk_vorticity_forces(topo.handle, _ptr(vel), _ptr(out), dt, inv_dx, scale, factor_scale,
                                       _ptr(src_vel), gravity, _stream(), int(asynchronous)))


def divergence(topo, vel, div, inv_dx, colour_split=False, asynchronous=False):
    _check(lib().hnsk_divergence(topo.handle, _ptr(vel), _ptr(div), inv_dx, int(colour_split), _stream(),
                                 int(asynchronous)))


def rbgs_sweep(topo, div_cs, p_cs, dx, colour, omega, asynchronous=False):
    _check(lib().hnsk_rbgs_sweep(topo.handle, _ptr(div_cs), _ptr(p_cs), dx, int(colour), omega, _stream(),
                                 int(asynchronous)))


def residual(topo, p_cs, div_cs, dx):
    r2, d2 = C.c_double(), C.c_double()
    _check(lib().hnsk_residual(topo.handle, _ptr(p_cs), _ptr(div_cs), dx, C.byref(r2), C.byref(d2), _stream()))
    return r2.value, d2.value


def pressure_solve(topo, div_cs, p_cs, dx, omega=1.0, iterations=0, solver="rbgs", tol=1e-5, max_iters=100000,
                   check_every=0, allow_unconverged=False):
    """Whole pressure solve from p = 0; returns hns_solve_stats. Raises unless converged, or with
    allow_unconverged returns the stats of a solve that hit max_iters first (p holds its result)."""
    from . import SOLVERS, hns_solve_options, hns_solve_stats
    o = hns_solve_options(C.sizeof(hns_solve_options), SOLVERS[solver], float(tol), int(max_iters), int(check_every))
    s = hns_solve_stats(C.sizeof(hns_solve_stats), 0, -1.0, 0.0)
    st = lib().hnsk_pressure_solve(topo.handle, _ptr(div_cs), _ptr(p_cs), dx, omega, int(iterations), C.byref(o),
                                   C.byref(s), _stream())
    if not (allow_unconverged and st == HNS_ERR_NOT_CONVERGED):
        _check(st)
    return s


def subtract_gradient(topo, vel, p_cs, out, inv_dx, sdf=None, asynchronous=False):
    _check(lib().hnsk_subtract_gradient(topo.handle, _ptr(vel), _ptr(p_cs), _ptr(out), inv_dx, _ptr(sdf), _stream(),
                                        int(asynchronous)))


def enforce_collision(topo, vel, sdf, voxel_size, asynchronous=False):
    _check(lib().hnsk_enforce_collision(topo.handle, _ptr(vel), _ptr(sdf), voxel_size, _stream(),
                                        int(asynchronous)))


def combustion_buoyancy(topo, fuel, waste, temp, flame, div, vel, vel_out, dt, temp_gain, expansion, ambient,
                        buoyancy, asynchronous=False):
    _check(lib().hnsk_combustion_buoyancy(topo.handle, _ptr(fuel), _ptr(waste), _ptr(temp), _ptr(flame), _ptr(div),
                                          _ptr(vel), _ptr(vel_out), dt, temp_gain, expansion, ambient, buoyancy,
                                          _stream(), int(asynchronous)))


def to_colour_split(topo, x, out, asynchronous=False):
    _check(lib().hnsk_to_colour_split(topo.handle, _ptr(x), _ptr(out), _stream(), int(asynchronous)))


def from_colour_split(topo, x, out, asynchronous=False):
    _check(lib().hnsk_from_colour_split(topo.handle, _ptr(x), _ptr(out), _stream(), int(asynchronous)))


def aos_to_soa(x, out, asynchronous=False):
    n = x.numel() // 3
    _check(lib().hnsk_aos_to_soa(n, _ptr(x), _ptr(out), _stream(), int(asynchronous)))


def soa_to_aos(x, out, asynchronous=False):
    n = x.numel() // 3
    _check(lib().hnsk_soa_to_aos(n, _ptr(x), _ptr(out), _stream(), int(asynchronous)))


def mg_apply(topo, r_cs, z_cs, graph=True, coarse_launches=False, unfused=False, asynchronous=False):
    """z = V(r): one multigrid V-cycle of the HNS_SOLVER_MGCG preconditioner (colour-split device
    vectors); graph replays the captured V-cycle, otherwise its kernels are launched one by one;
    coarse_launches runs the coarsest level as one launch per colour sweep, unfused level 0 as
    separate sweeps, restriction and prolongation (same bits)."""
    flags = (1 if graph else 0) | (2 if coarse_launches else 0) | (4 if unfused else 0)
    _check(lib().hnsk_mg_apply(topo.handle, _ptr(r_cs), _ptr(z_cs), flags, _stream()))
    if not asynchronous:
        import torch
        torch.cuda.synchronize()
