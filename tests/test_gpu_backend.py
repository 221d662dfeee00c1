"""The C++ plugin interface (IComputeBackend over libmxhip.so, built into
mandelbrot-explorer_amd/_build/backend_driver) on the GPU: request validation,
cancellation, strips, and a golden frame rendered through it."""
import os
import subprocess

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRV = os.path.join(ROOT, "mandelbrot-explorer_amd", "_build", "backend_driver")


def test_backend_checks():
    r = subprocess.run([DRV, "checks"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr


def test_backend_renders_golden_frame(tmp_path):
    cre, cim, scale, W, H, mxit, cm = cases.FRAMES["extref875"]
    out = tmp_path / "f.f32"
    r = subprocess.run([DRV, "frame", cre, cim, scale, str(W), str(H), str(mxit), str(cm), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    it = np.fromfile(out, np.float32).reshape(H, W)
    g = cases.golden("minibrot875_extref.f32", H, W)
    assert np.array_equal(it.view(np.int32), g.view(np.int32))
