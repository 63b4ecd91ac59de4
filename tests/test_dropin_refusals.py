"""The Fortran drop-in refuses, before touching the GPU, every input that the
reference reads on the per-step path but the device does not implement
(FWGPU_INIT), and the host-side per-step diagnostics that would read stale
MODULE GLOBAL arrays (FWGPU_BEGIN).  Runs on the CPU: the refusals come
after the reference's own READ_INPUT and INITIALIZATION and before any
library call that needs a device."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from common import REPO

BIN = os.path.join(REPO, "oracle", "_ref", "funwave_gpu_sp")
pytestmark = pytest.mark.skipif(not os.path.exists(BIN),
                                reason="drop-in program not built (make -C oracle dropin)")

REFUSED = {
    "HIGH_ORDER": ({"HIGH_ORDER": "THIRD"}, "HIGH_ORDER"),
    "CONSTRUCTION": ({"CONSTRUCTION": "NONE"}, "CONSTRUCTION"),
    # (the reference itself stops on WAVEMAKER_VIS with VISCOSITY_BREAKING)
    "WAVEMAKER_VIS": ({"WAVEMAKER_VIS": True, "VISCOSITY_BREAKING": False, "visbrk": 0.1,
                       "WAVEMAKER_visbrk": 0.1}, "WAVEMAKER_VIS"),
    "WaveMakerCurrentBalance": ({"WaveMakerCd": 0.01}, "WaveMakerCurrentBalance"),
    "BREAKWATER": ({"BREAKWATER_FILE": "breakwater.txt"}, "BREAKWATER"),
    "WAVEMAKER": ({"WAVEMAKER": "LEF_SOL", "AMP": 0.1, "DEP": 2.0, "LAGTIME": 1.0}, "WAVEMAKER"),
    "STEADY_TIME": ({"STEADY_TIME": 0.0}, "STEADY_TIME"),
    "Hmax": ({"Hmax": True}, "MAX_MIN_PROPERTY"),
}


@pytest.mark.parametrize("key", list(REFUSED))
def test_unsupported_input_is_refused(key):
    import cases
    extra, word = REFUSED[key]
    with tempfile.TemporaryDirectory() as d:
        c = cases.tiny2d()
        c.params.update(extra)
        c.params["TOTAL_TIME"] = 1.0
        c.write(d)
        if "BREAKWATER_FILE" in extra:
            np.savetxt(os.path.join(d, "breakwater.txt"), np.zeros((c.params["Nglob"], c.params["Mglob"])))
        r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; exec {BIN} input.txt"], cwd=d,
                           capture_output=True, text=True, timeout=120, env=dict(os.environ, HIP_VISIBLE_DEVICES=""))
    out = r.stdout + r.stderr
    assert r.returncode == 2, out[-1500:]
    assert "FWGPU" in out and word in out, out[-1500:]


def test_supported_input_passes_the_checks():
    """The same case without those inputs gets past the checks (it then
    fails only where a device is needed, when there is none)."""
    import cases
    with tempfile.TemporaryDirectory() as d:
        c = cases.tiny2d()
        c.params["TOTAL_TIME"] = 1.0
        c.write(d)
        r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; exec {BIN} input.txt"], cwd=d,
                           capture_output=True, text=True, timeout=120, env=dict(os.environ, HIP_VISIBLE_DEVICES=""))
    assert "unsupported on the device path" not in r.stdout
