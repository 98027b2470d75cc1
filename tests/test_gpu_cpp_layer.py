"""The MAST-style C++ layer (mast-multiphysics_amd/host/mast_gpu.hpp) drives the
reference's call sequence (solve, ComplianceOutput, adjoint_solve, adjoint
sensitivities) through the C ABI and agrees with the fused design evaluation."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mast-multiphysics_amd", "example_compliance")

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nx", [6, 16])
def test_mast_call_sequence(nx):
    print(f"[C++ layer] nx={nx}", flush=True)
    assert os.path.exists(EXE), "example_compliance is not built"
    p = subprocess.run([EXE, str(nx)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr + p.stdout
    out = json.loads(p.stdout.strip().splitlines()[-1])
    assert out["compliance_rel_diff"] < 1e-8
    assert out["sens_max_rel_diff"] < 1e-6
