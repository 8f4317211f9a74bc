"""HDF5 field I/O (zmodel/utils.cu:103-281) on the GPU path.

* A device field in the packed layout is stored with store_host_hdf5, read
  back with init_host_hdf5 into a new device tensor, and must come back bit
  for bit; transforming the read-back field gives the bits of transforming
  the original.  A second distribution of the same global field reads the
  file into its own packed layout, which the device converter maps back to
  the same natural field.
* The C++ facade example (stencilfft_amd/examples/hdf5_example.cpp) stores the
  ramp, loads it, runs Laplacian -> forward -> inverse on the GPU and stores
  and reloads the result; it checks the file round trips bit for bit and the
  result against the ramp's periodic Laplacian.
"""
import os
import subprocess

import pytest

from conftest import ROOT, has_gpu

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a GPU")]

EXAMPLE = os.path.join(ROOT, "stencilfft_amd", "examples", "zmodel_hdf5_example")


def _h5():
    from stencilfft_amd import hdf5
    if not os.path.exists(hdf5.HDF5_LIB_PATH):
        pytest.skip("libstencilfft_hdf5.so was not built (no HDF5 C library)")
    return hdf5


@pytest.mark.parametrize("dims,n,p,b", [(2, 1024, 4, 16), (3, 64, 2, 4)])
def test_device_field_file_round_trip(tmp_path, dims, n, p, b):
    import torch

    from stencilfft_amd import Plan
    from stencilfft_amd.fields import uniform_fields
    h5 = _h5()
    print(f"[gpu] hdf5 device round trip {dims}D N={n} P={p} B={b}", flush=True)
    pl = Plan(dims, n, p, b, precision="fp64")
    x = uniform_fields(pl, seed=n + dims)
    fn = str(tmp_path / "field.h5")
    h5.store_host_hdf5(pl, x, fn, "u")
    back = torch.from_numpy(h5.init_host_hdf5(pl, fn, "u")).cuda()
    assert torch.equal(back.view(torch.float64), x.view(torch.float64))
    assert torch.equal(pl.forward(back).view(torch.float64), pl.forward(x).view(torch.float64))
    # the same global field read into another distribution (P, B swapped roles)
    other = Plan(dims, n, p // 2 if p > 2 else p, b * 2 if p > 2 else b // 2, precision="fp64")
    y = torch.from_numpy(h5.init_host_hdf5(other, fn, "u")).cuda()
    assert torch.equal(other.packed_to_natural(y), pl.packed_to_natural(x))


@pytest.mark.parametrize("args", [["2", "256", "4", "8"], ["3", "32", "2", "4"]])
def test_cpp_facade_hdf5_gpu_example(tmp_path, args):
    _h5()
    print(f"[gpu] zmodel_hdf5_example {' '.join(args)}", flush=True)
    assert os.path.exists(EXAMPLE), "build the library first (python -m stencilfft_amd.build)"
    res = subprocess.run([EXAMPLE, str(tmp_path / "ramp.h5"), *args], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0 and "OK" in res.stdout
