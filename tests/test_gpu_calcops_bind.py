"""libCalcOps_bind.so as the reference's models load it: a child process that never imports the
Python operator registration reaches the library through <src>/op/build/lib/libCalcOps_bind.so
(src/model/dp_dp.py:13-20, src/model/nep_net.py:17-24) and runs all ten CalcOps_cuda operators
against the FP64 oracle:
  * calculate_maxneigh / calculate_neighbor: bit-exact counts, lists and [r, x, y, z];
  * calculateNepFeat / calculateNepMbFeat / calculate_descriptor: features to 1e-11 relative;
  * a NEP model chained on the operators as the reference's GPU branch chains them
    (calculateNepForce with the centre term by matmul, calculateNepVirial): E, F, V to 1e-10
    relative in float64, force RMSE 1e-4 eV/A in float32;
  * the force-loss gradient to c2, c3 and the fitting parameters through the operators' double
    backward: 1e-9 relative to the oracle's autograd;
  * calculateForce / calculateVirial / calculateCompress: 1e-12 relative to the oracle's ops."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(os.path.dirname(HERE), "matpl_amd", "libCalcOps_bind.so")


def test_ten_operators_through_the_reference_load_path(tmp_path):
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} is not built")
    print("case libCalcOps_bind.so, ten operators (child process)", flush=True)
    r = subprocess.run([sys.executable, "-u", os.path.join(HERE, "calcops_bind_child.py"), str(tmp_path)],
                       capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    e = json.loads(r.stdout.strip().splitlines()[-1])
    print(e, flush=True)
    assert not e["python_ops_imported"]
    assert e["nl_exact"] and e["maxneigh_exact"], e
    assert e["dq"] < 1e-11 and e["d_descriptor"] < 1e-11, e
    assert e["dE"] < 1e-10 and e["dF"] < 1e-10 and e["dV"] < 1e-10, e
    assert e["d_loss"] < 1e-10 and e["d_grad"] < 1e-9, e
    assert e["dF_fp32_rmse"] < 1e-4, e
    assert e["dp_force"] < 1e-12 and e["dp_virial"] < 1e-12 and e["dp_compress"] < 1e-12, e
