"""CPU checks of the C ABI: the library loads, exports every symbol include/hns.h declares, and
answers the host-side error paths without a device (null contexts, empty payloads, bad
arguments). Also builds the header-only C++ adapter (include/hns.hpp) into a program shaped like
the reference's SOP caller and runs its error paths."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import hnanosolver_amd as hns
from hnanosolver_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hns.h")


def declared():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:hns_status|const char\*|int32_t|int64_t)\s+(hnsk?_\w+)\s*\(", src, re.M)))


def test_every_declared_symbol_is_exported_and_bound():
    names = declared()
    assert len(names) >= 30
    lib = C.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.EXPORTS) == names
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (hnsk?_\w+)", out))
    assert exported == set(names)


def test_abi_version_and_param_layout():
    assert hns.lib().hns_abi_version() == 1
    assert C.sizeof(hns.hns_params) == 72
    offs = [getattr(hns.hns_params, f).offset for f, _ in hns.hns_params._fields_]
    assert offs == list(range(0, 72, 4))


def test_null_context_paths():
    lib = hns.lib()
    p = hns.make_params()
    assert lib.hns_advance(None, 10, 0.1, C.byref(p)) == -3
    assert lib.hns_upload_sources(None, None) == -3
    assert lib.hns_download_for_output(None, None) == -3
    assert b"null context" in lib.hns_last_error()
    n = C.c_size_t(99)
    assert lib.hns_get_active_mask(None, None, 0, C.byref(n)) == 0 and n.value == 0
    h = C.c_void_p()
    assert lib.hns_destroy_context(C.byref(h)) == 0
    assert lib.hns_destroy_context(None) == 0
    assert lib.hns_set_timing(None, 1) == -3


def test_empty_and_bad_payloads():
    lib = hns.lib()
    d = hns.GridIndexedData(np.zeros((0, 3), np.int32))
    with pytest.raises(hns.HNSError) as e:
        hns.create_index_grid(d, 0.1)
    assert e.value.status == -2
    with pytest.raises(hns.HNSError) as e:
        hns.ensure_context(None, d, hns.Topology(None), 0.1)
    assert e.value.status == -2
    h = C.c_void_p()
    assert lib.hns_create_index_grid(None, 0.1, C.byref(h)) == -1
    assert lib.hns_leaf_order(None, 3, None) == -1
    assert lib.hns_leaf_order(None, 0, None) == 0


def test_no_device_is_reported_not_crashed():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a HIP device is visible; the no-device path is covered on CPU-only hosts")
    c = np.zeros((512, 3), np.int32)
    c[:] = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij"), -1).reshape(512, 3)
    d = hns.GridIndexedData(c)
    d.add_value_block("velocity", np.zeros((512, 3), np.float32))
    with pytest.raises(hns.HNSError) as e:
        hns.create_index_grid(d, 0.1)
    assert e.value.status == -10
    with pytest.raises(hns.HNSError) as e:
        hns.advect_index_grid_velocity(d, 0.1, 0.1)
    assert e.value.status == -10


def test_payload_validation():
    d = hns.GridIndexedData(np.zeros((512, 3), np.int32))
    with pytest.raises(ValueError):
        d.add_value_block("x", np.zeros(511, np.float32))
    d.add_value_block("x", np.zeros(512, np.float32))
    with pytest.raises(ValueError):
        d.add_value_block("x", np.zeros(512, np.float32))


def build_adapter(out):
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "adapter_cook.cpp"), "-o", out, "-L", os.path.dirname(_lib.LIB_PATH),
           "-lhns", "-Wl,-rpath," + os.path.dirname(_lib.LIB_PATH)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def test_cpp_adapter_compiles_and_error_paths(tmp_path):
    exe = build_adapter(str(tmp_path / "adapter_cook"))
    r = subprocess.run([exe, "errors"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "errors: ok" in r.stdout
