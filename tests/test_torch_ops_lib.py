"""The C++ operator library (libCalcOps_bind.so) loads without a GPU and registers the ten
CalcOps_cuda operators of the reference (src/op/src/CalcOps_bind.cpp:6-17) with the schemas the
reference's m.def(name, fn) infers from its C++ signatures (src/op/include/CalcOps.h:35-318), plus
matpl_amd::env_mat. Checked in a child process: this one may hold the Python registration."""
import json
import os
import subprocess
import sys

from matpl_amd import dp_script

T = "Tensor"
# argument types in order, as inferred from CalcOps.h (double -> float, int64_t -> int)
WANT = {
    "calculateForce": [T] * 5,
    "calculateVirial": [T] * 5,
    "calculateCompress": [T] * 2,
    "calculateNepFeat": [T] * 5 + ["float", "int"],
    "calculateNepMbFeat": [T] * 5 + ["int", "int", "int", "int", "float"],
    "calculateNepForce": [T] * 4,
    "calculateNepVirial": [T] * 5,
    "calculate_maxneigh": [T] * 5 + ["float", "float", "int", T, "bool"],
    "calculate_neighbor": [T] * 7 + ["float", "float", "int", "int", "bool"],
    "calculate_descriptor": [T] * 5 + ["float", "float", "int", "int", "int", "int"],
}


def test_ops_library_schemas():
    assert os.path.exists(dp_script.OPS_LIB), f"{dp_script.OPS_LIB} is not built"
    assert os.path.basename(dp_script.OPS_LIB) == "libCalcOps_bind.so"
    code = ("import torch, json; torch.ops.load_library(%r); ns = torch.ops.CalcOps_cuda; "
            "s = {n: [str(a.type) for a in getattr(ns, n).default._schema.arguments] + "
            "[str(r.type) for r in getattr(ns, n).default._schema.returns] for n in %r}; "
            "s['env_mat'] = str(torch.ops.matpl_amd.env_mat.default._schema); print(json.dumps(s))"
            % (dp_script.OPS_LIB, sorted(WANT)))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    s = json.loads(r.stdout.strip().splitlines()[-1])
    for name, args in WANT.items():
        assert s[name] == args + ["List[Tensor]"], (name, s[name])
    assert s["env_mat"].startswith("matpl_amd::env_mat(Tensor image_dr")
