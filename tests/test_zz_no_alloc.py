"""hnsk_residual_noalloc (the allocation-free residual) against hnsk_residual through the C ABI:
bitwise the same r^2 and ||div||^2 on two calls in a row, the second on a smaller topology (the
shared static scratch keeps the first call's partial sums), no change in free device memory
around either call (the scratch's address is looked up when the topology is built), and the size
limit rejected with a clear error."""
import ctypes as C

import numpy as np
import pytest

import hnanosolver_amd as hns
from hnanosolver_amd import kernels as K
from hnanosolver_amd import scenes

from _helpers import canonical_coords, ragged_origins

pytestmark = pytest.mark.gpu


def _case(name):
    print("CASE", name, flush=True)


def _pair(torch, n, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return torch.randn(n, device="cuda", generator=g), torch.randn(n, device="cuda", generator=g)


def _noalloc(torch, topo, p, d, dx, out):
    st = hns.lib().hnsk_residual_noalloc(topo.handle, C.c_void_p(p.data_ptr()), C.c_void_p(d.data_ptr()), dx,
                                         C.c_void_p(out.data_ptr()), C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == 0, hns.lib().hns_last_error()
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.timeout(120, method="thread")
def test_residual_noalloc_matches_residual_twice_second_smaller(hip):
    torch = hip
    big = canonical_coords(ragged_origins(7))
    small = scenes.smoke64()
    assert small.n < len(big)
    cases = [("ragged7", big, 0.05, 11), ("smoke64", small.coords, 0.1, 12)]
    for name, coords, dx, seed in cases:
        _case("residual_noalloc %s N=%d" % (name, len(coords)))
        topo = hns.create_index_grid(hns.GridIndexedData(coords), dx)
        p, d = _pair(torch, len(coords), seed)
        r2, d2 = K.residual(topo, p, d, dx)
        # the output exists before the first reading (torch takes small tensors from 2 MiB segments)
        out = torch.full((2,), -1.0, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        got = _noalloc(torch, topo, p, d, dx, out)
        free1 = torch.cuda.mem_get_info()[0]
        print("  hnsk_residual %.17g %.17g | noalloc %.17g %.17g | free memory %d -> %d" % (
            r2, d2, got[0], got[1], free0, free1))
        assert got[0] == r2 and got[1] == d2
        assert free1 == free0


@pytest.mark.timeout(120, method="thread")
def test_residual_noalloc_rejects_topologies_above_the_limit(hip):
    torch = hip
    L = hns.lib()
    limit = 131072  # HNS_NOALLOC_MAX_LEAVES
    _case("residual_noalloc over the limit: %d leaves" % (limit + 1))
    origins = np.zeros((limit + 1, 3), np.int32)
    origins[:, 0] = 8 * np.arange(limit + 1, dtype=np.int32)
    loc = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij"), -1).reshape(512, 3)
    coords = (origins[:, None, :] + loc[None, :, :].astype(np.int32)).reshape(-1, 3)
    topo = hns.create_index_grid(hns.GridIndexedData(coords), 0.1)
    n = len(coords)
    p = torch.zeros(n, device="cuda")
    out = torch.full((2,), -1.0, dtype=torch.float64, device="cuda")
    st = L.hnsk_residual_noalloc(topo.handle, C.c_void_p(p.data_ptr()), C.c_void_p(p.data_ptr()), 0.1,
                                 C.c_void_p(out.data_ptr()), C.c_void_p(torch.cuda.current_stream().cuda_stream))
    msg = L.hns_last_error().decode()
    print("  status %d: %s" % (st, msg))
    assert st == -1 and "HNS_NOALLOC_MAX_LEAVES" in msg
    torch.cuda.synchronize()
    assert (out.cpu().numpy() == -1.0).all()
