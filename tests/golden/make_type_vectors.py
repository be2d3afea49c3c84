"""Writes tests/golden/type_vectors.json: the normalized physical-property type vectors of every
element the reference's table covers, for its five default properties (src/user/nn_feature_type.py:35),
computed from the table in utils/atom_type_emb_dict.py the way its get_normalized_data does
((value - min) / (max - min) over the property's non-NaN entries; atomic numbers over the table's
keys 1..109), with null where that function refuses the element (NaN or no entry). The table is
read as text (parsed, not imported). Run in a tree where the reference is at /root/reference."""
import ast
import json
import math
import os

SRC = "/root/reference/utils/atom_type_emb_dict.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "type_vectors.json")
PROPS = ["atomic_number", "atom_radius", "atom_mass", "electron_affin", "pauling"]


def _table():
    tree = ast.parse(open(SRC).read())
    for node in tree.body:
        if isinstance(node, ast.Assign) and getattr(node.targets[0], "id", "") == "physical":
            out = {}
            for k, v in zip(node.value.keys, node.value.values):
                if not isinstance(v, ast.Dict):
                    continue
                col = {}
                for kk, vv in zip(v.keys, v.values):
                    if isinstance(vv, ast.Call):                 # float("NaN")
                        col[kk.value] = float("nan")
                    elif isinstance(vv, ast.UnaryOp):
                        col[kk.value] = -vv.operand.value
                    else:
                        col[kk.value] = vv.value
                out[k.value] = col
            return out
    raise RuntimeError("no 'physical' table")


def main():
    t = _table()
    names = t["atomic_number"]                                   # {Z: symbol}
    scale = {}
    for p in PROPS:
        vals = list(names.keys()) if p == "atomic_number" else [v for v in t[p].values() if not math.isnan(v)]
        scale[p] = (min(vals), max(vals))
    vec = {}
    for z, sym in names.items():
        row = []
        for p in PROPS:
            v = z if p == "atomic_number" else t[p].get(sym)
            row.append(None if v is None or math.isnan(v) else (v - scale[p][0]) / (scale[p][1] - scale[p][0]))
        vec[str(z)] = row
    json.dump({"properties": PROPS, "symbols": {str(z): s for z, s in names.items()}, "vectors": vec}, open(OUT, "w"),
              indent=0)


if __name__ == "__main__":
    main()
