"""The C++ drop-in layer (salty_amd/host/s3d_gpu.h) end to end: the Cornell
box built with Mesh::Create / MaterialFactory / Scene / PathTracer::Run renders
the same image as the oracle at the same seed and pass count."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from salty_amd import scenes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "salty_amd", "host", "s3d_cornell")


def test_cpp_host_cornell_matches_oracle(tmp_path):
    w = h = 64
    passes = 4
    ppm = tmp_path / "out.ppm"
    sums = tmp_path / "sum.bin"
    out = subprocess.run([EXE, str(w), str(h), str(passes), str(ppm), str(sums)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    img = np.fromfile(sums, dtype=np.float32).reshape(h, w, 3)
    ref, _ = O.OracleScene(scenes.cornell_box()).render(w, h, spp=passes, max_depth=8, seed=0)
    rel = np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2))
    assert rel < 1e-4, rel
    data = ppm.read_bytes()
    assert data.startswith(b"P6\n64 64\n255\n") and len(data) == len(b"P6\n64 64\n255\n") + w * h * 3
