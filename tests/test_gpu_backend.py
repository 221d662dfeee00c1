"""The C++ plugin interface (IComputeBackend over libmxhip.so, built into
mandelbrot-explorer_amd/_build/backend_driver) on the GPU: request validation,
cancellation before and during a frame, strips against the full frame, the CPU
fallback for requests off this path, and golden frames rendered through it
(verify.cpp:9668-10069)."""
import os
import subprocess

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRV = os.path.join(ROOT, "mandelbrot-explorer_amd", "_build", "backend_driver")


def run(*args, timeout=120):
    print("[backend]", " ".join(a[:24] for a in args), flush=True)
    r = subprocess.run([DRV, *args], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_backend_checks():
    run("checks")


def test_backend_cancel_mid_frame():
    out = run("cancel", cases.LOC["deep1_x"], cases.LOC["deep1_y"])
    assert "cancel ok" in out


def frame(tmp_path, name, W, H, mxit, cm, sub=1, cre=None, cim=None, scale=None):
    out = tmp_path / f"{name}.f32"
    run("frame", cre, cim, scale, str(W), str(H), str(mxit), str(cm), str(sub), str(out))
    return np.fromfile(out, np.float32).reshape(H * sub, W * sub)


def test_backend_renders_golden_frame(tmp_path):
    cre, cim, scale, W, H, mxit, cm = cases.FRAMES["extref875"]
    it = frame(tmp_path, "extref", W, H, mxit, cm, cre=cre, cim=cim, scale=scale)
    g = cases.golden("minibrot875_extref.f32", H, W)
    assert np.array_equal(it.view(np.int32), g.view(np.int32))


def test_backend_adaptive_supersampling_golden(tmp_path):
    """A GUI request with sub 5 and SUPER_SAMPLING through the plugin gives the
    reference's adaptive golden (runAdaptiveGuiCase, verify.cpp:307-388)."""
    it = frame(tmp_path, "gui875", 4, 4, 300000, 1, sub=5, cre=cases.LOC["deep1_x"], cim=cases.LOC["deep1_y"],
               scale=cases.pow10(875))
    g = cases.golden("minibrot875_gui_ss5.f32", 20, 20)
    assert np.array_equal(it.view(np.int32), g.view(np.int32))
