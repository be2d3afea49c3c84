"""The TorchScript drop-in for the LAMMPS pair style (pair_matpl.cpp:864-1000 loads a scripted model
and calls torch.ops.CalcOps_cuda.* from C++): a DP model scripted against the C++ operator library
alone, saved and reloaded, reproduces the FP64 oracle on a local + ghost system. Runs in a child
process, since this one has the Python CalcOps_cuda operators registered."""
import json
import os
import subprocess
import sys

import pytest

from matpl_amd import dp_script

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_scripted_dp_matches_oracle(tmp_path):
    if not os.path.exists(dp_script.OPS_LIB):
        pytest.fail(f"{dp_script.OPS_LIB} is not built")
    print("case scripted DP hfo2 2x2x2 + ghosts (child process)", flush=True)
    r = subprocess.run([sys.executable, "-u", os.path.join(HERE, "scripted_dp_child.py"), str(tmp_path / "dp.pt")],
                       capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    err = json.loads(r.stdout.strip().splitlines()[-1])
    assert err["dE"] < 1e-10 and err["dEi"] < 1e-10, err
    assert err["dF"] < 1e-9 and err["dV"] < 1e-8, err

