"""Writes aat_goldens.json: the finite-difference AAT arrays held by the
reference's own tests (apyib/tests/test_011_AAT.py, test_012_AAT_SO.py,
test_013_AAT_parallel.py) with each test's parameters.

The test files are parsed as text with `ast` (literal lists and dicts only);
nothing from the reference is imported or executed.  Commented-out tests are
not seen by `ast` and are therefore not recorded.

    python tests/golden/make_aat_goldens.py /root/reference/apyib/tests
"""
import ast
import json
import os
import sys

FILES = ("test_011_AAT.py", "test_012_AAT_SO.py", "test_013_AAT_parallel.py")


def _literal(node):
    return ast.literal_eval(node)


def _array(node):
    """np.array([...]) -> nested list."""
    if isinstance(node, ast.Call) and getattr(node.func, "attr", "") == "array":
        return _literal(node.args[0])
    return None


def _sum_names(node, arrays):
    if isinstance(node, ast.Name):
        return arrays[node.id]
    if isinstance(node, ast.BinOp) and isinstance(node.op, ast.Add):
        a, b = _sum_names(node.left, arrays), _sum_names(node.right, arrays)
        return [[x + y for x, y in zip(ra, rb)] for ra, rb in zip(a, b)]
    raise ValueError("unsupported aat_ref expression")


def _parameters(node):
    out = {}
    for k, v in zip(node.keys, node.values):
        key = _literal(k)
        if isinstance(v, ast.Subscript):           # moldict["..."]
            out[key] = {"moldict": _literal(v.slice)}
        else:
            out[key] = _literal(v)
    return out


def _calls(fn):
    for node in ast.walk(fn):
        if isinstance(node, ast.Call):
            name = getattr(node.func, "attr", getattr(node.func, "id", ""))
            yield name, node


def parse_test(fn, fname):
    arrays, params, aat = {}, None, None
    for stmt in ast.walk(fn):
        if isinstance(stmt, ast.Assign) and len(stmt.targets) == 1 and isinstance(stmt.targets[0], ast.Name):
            tgt, val = stmt.targets[0].id, stmt.value
            if tgt == "parameters" and isinstance(val, ast.Dict):
                params = _parameters(val)
            elif _array(val) is not None:
                arrays[tgt] = _array(val)
            elif tgt == "aat_ref":
                aat = _sum_names(val, arrays)
    if aat is None:
        aat = arrays.get("aat_ref")
    dR = dH = None
    normalization = "full"
    form = None
    for name, call in _calls(fn):
        if name in ("compute_parallel_aats", "compute_AAT"):
            nums = [a for a in call.args if isinstance(a, ast.Constant) and isinstance(a.value, float)]
            dR, dH = nums[0].value, nums[1].value
        if name in ("compute_parallel_aats", "compute_spatial_aats", "compute_SO_aats"):
            for kw in call.keywords:
                if kw.arg == "normalization":
                    normalization = _literal(kw.value)
            for a in call.args:
                if isinstance(a, ast.Constant) and isinstance(a.value, str):
                    normalization = a.value
        if name == "compute_spatial_aats":
            form = "spatial"
        if name == "compute_SO_aats":
            form = "SO"
    tol = None
    for node in ast.walk(fn):
        if isinstance(node, ast.Assert) and isinstance(node.test, ast.Compare):
            tol = _literal(node.test.comparators[0])
    skipped = any("skip" in ast.unparse(d) for d in fn.decorator_list)
    if form is None:                               # compute_parallel_aats picks by method
        form = "SO" if params["method"].endswith("_SO") else "spatial"
    return {"file": fname, "line": fn.lineno, "test": fn.name, "parameters": params, "dR": dR, "dH": dH,
            "normalization": normalization, "form": form, "tolerance": tol, "skipped_in_reference": skipped,
            "components": {k: v for k, v in arrays.items() if k != "aat_ref"}, "aat": aat}


def main(test_dir):
    out = []
    for fname in FILES:
        with open(os.path.join(test_dir, fname)) as f:
            tree = ast.parse(f.read())
        for node in tree.body:
            if isinstance(node, ast.FunctionDef) and node.name.startswith("test_"):
                out.append(parse_test(node, fname))
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "aat_goldens.json")
    with open(dst, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"wrote {len(out)} AAT goldens to {dst}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/apyib/tests")
