"""The C++ operator library for TorchScript hosts (libcalcops_torch.so) loads without a GPU and
registers the reference's CalcOps_cuda DP operator names and schemas (CalcOps_bind.cpp:6-17) plus
matpl_amd::env_mat. Checked in a child process: this one registers the Python CalcOps_cuda ops."""
import json
import os
import subprocess
import sys

from matpl_amd import dp_script


def test_ops_library_schemas():
    assert os.path.exists(dp_script.OPS_LIB), f"{dp_script.OPS_LIB} is not built"
    code = ("import torch, json; torch.ops.load_library(%r); "
            "print(json.dumps([str(torch.ops.CalcOps_cuda.calculateForce.default._schema), "
            "str(torch.ops.CalcOps_cuda.calculateVirial.default._schema), "
            "str(torch.ops.CalcOps_cuda.calculateCompress.default._schema), "
            "str(torch.ops.matpl_amd.env_mat.default._schema)]))" % dp_script.OPS_LIB)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    s = json.loads(r.stdout.strip().splitlines()[-1])
    assert s[0].startswith("CalcOps_cuda::calculateForce(Tensor") and s[0].count("Tensor") == 6
    assert s[1].startswith("CalcOps_cuda::calculateVirial(Tensor") and s[1].count("Tensor") == 6
    assert s[2].startswith("CalcOps_cuda::calculateCompress(Tensor")
    assert s[3].startswith("matpl_amd::env_mat(Tensor image_dr")
