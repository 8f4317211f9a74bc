"""The C++ drop-in example (zmodel/main.cu example() over the zmodel.hpp facade)
runs on the GPU and reproduces the host-computed Laplacian after the FFT round trip."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "stencilfft_amd", "examples", "zmodel_example")


@pytest.mark.parametrize("args", [[], ["64", "4", "2"], ["2048", "4", "16"]])
def test_example_round_trip(args):
    print(f"[gpu] zmodel_example {' '.join(args)}", flush=True)
    assert os.path.exists(BIN), "build the library first (python -m stencilfft_amd.build)"
    res = subprocess.run([BIN, *args], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0
    assert "OK" in res.stdout
