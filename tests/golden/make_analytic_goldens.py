"""Writes analytic_aat_goldens.json: the analytic AAT arrays held by the
reference's tests (apyib/tests/test_018_RHF_Analytic_AAT.py,
test_019_MP2_Analytic_AAT.py, test_020_CID_Analytic_AAT.py,
test_021_CISD_Analytic_AAT.py) with each test's parameters, geometry
string, the compute_*_AATs call (method, normalization, orbitals) and the
assert tolerance.

The test files are parsed as text with `ast` (literals only); nothing from
the reference is imported or executed.

    python tests/golden/make_analytic_goldens.py /root/reference/apyib/tests
"""
import ast
import json
import os
import sys

FILES = ("test_018_RHF_Analytic_AAT.py", "test_019_MP2_Analytic_AAT.py", "test_020_CID_Analytic_AAT.py",
         "test_021_CISD_Analytic_AAT.py")


def _array(node):
    if isinstance(node, ast.Call) and getattr(node.func, "attr", "") == "array":
        return ast.literal_eval(node.args[0])
    return None


def parse_test(fn, fname):
    strings, params, aat, call = {}, None, None, None
    for stmt in ast.walk(fn):
        if isinstance(stmt, ast.Assign) and len(stmt.targets) == 1 and isinstance(stmt.targets[0], ast.Name):
            tgt, val = stmt.targets[0].id, stmt.value
            if isinstance(val, ast.Constant) and isinstance(val.value, str):
                strings[tgt] = val.value
            elif tgt == "parameters" and isinstance(val, ast.Dict):
                params = {}
                for k, v in zip(val.keys, val.values):
                    key = ast.literal_eval(k)
                    if isinstance(v, ast.Name):
                        params[key] = {"geometry": strings[v.id]}
                    elif isinstance(v, ast.Subscript):
                        params[key] = {"moldict": ast.literal_eval(v.slice)}
                    else:
                        params[key] = ast.literal_eval(v)
            elif tgt == "aat_ref" and _array(val) is not None:
                aat = _array(val)
            elif tgt == "aat" and isinstance(val, ast.Call):
                name = getattr(val.func, "attr", "")
                kw = {k.arg: ast.literal_eval(k.value) for k in val.keywords}
                call = {"function": name, **kw}
    tol = None
    for node in ast.walk(fn):
        if isinstance(node, ast.Assert) and isinstance(node.test, ast.Compare):
            tol = ast.literal_eval(node.test.comparators[0])
    skipped = any("skip" in ast.unparse(d) for d in fn.decorator_list)
    return {"file": fname, "line": fn.lineno, "test": fn.name, "parameters": params, "call": call,
            "tolerance": tol, "skipped_in_reference": skipped, "aat": aat}


def main(test_dir):
    out = []
    for fname in FILES:
        with open(os.path.join(test_dir, fname)) as f:
            tree = ast.parse(f.read())
        for node in tree.body:
            if isinstance(node, ast.FunctionDef) and node.name.startswith("test_"):
                out.append(parse_test(node, fname))
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "analytic_aat_goldens.json")
    with open(dst, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"wrote {len(out)} analytic AAT goldens to {dst}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/apyib/tests")
