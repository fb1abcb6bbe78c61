"""The torch.autograd.Function wrapper (raptor_amd/autograd.py): gradient of a
weighted image sum against central differences (step 1e-3) of the FP64
oracle.  Runs in a fresh process so torch loads before the library."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_autograd_gradient_matches_oracle_differences():
    r = subprocess.run([sys.executable, os.path.join(HERE, "autograd_check.py")],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["smooth_frac"] > 0.85
    assert out["params_restored"] and out["integrator_kept"]
    assert out["loss"] == pytest.approx(out["loss_oracle"], rel=1e-4)
    for g, f in zip(out["grad"], out["fd"]):
        assert g == pytest.approx(f, rel=2e-3)
