"""Writes curvilinear_kats.json: the known-answer arrays of the reference's
polar, cylindrical and spherical differentiator tests
(tests/operators/fdm/test_numerical_differentiator.py:670-2457), for the
unconstrained gradient, hessian, divergence, curl, laplacian and
vector_laplacian cases.

The reference test file is parsed as text with ``ast``; only literal
numbers, lists, arithmetic, ``np.pi``, ``np.array`` and ``Mesh(...)`` /
``CoordinateSystem.X`` expressions are evaluated, by the small evaluator
below.  Nothing from the reference is imported or executed.

Usage: python tests/golden/make_curvilinear_kats.py [path to the PararealML
source tree, default /root/reference]
"""

import ast
import json
import math
import re
import sys
from pathlib import Path

OPS = ("gradient", "hessian", "divergence", "curl", "laplacian", "vector_laplacian")
NAME = re.compile(r"^test_3pcfdm_(?:mixed_)?(polar|cylindrical|spherical)_(\w+)$")


class Skip(Exception):
    pass


def value(node, env):
    if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
        return node.value
    if isinstance(node, (ast.List, ast.Tuple)):
        return [value(e, env) for e in node.elts]
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
        v = value(node.operand, env)
        return -v if isinstance(node.op, ast.USub) else v
    if isinstance(node, ast.BinOp):
        a, b = value(node.left, env), value(node.right, env)
        ops = {ast.Add: lambda: a + b, ast.Sub: lambda: a - b, ast.Mult: lambda: a * b,
               ast.Div: lambda: a / b, ast.Pow: lambda: a ** b}
        if type(node.op) not in ops:
            raise Skip("operator")
        return ops[type(node.op)]()
    if isinstance(node, ast.Attribute) and isinstance(node.value, ast.Name):
        if node.value.id == "np" and node.attr == "pi":
            return math.pi
        if node.value.id == "CoordinateSystem":
            return node.attr
        raise Skip(f"attribute {node.value.id}.{node.attr}")
    if isinstance(node, ast.Name):
        if node.id in env:
            return env[node.id]
        raise Skip(f"name {node.id}")
    if isinstance(node, ast.Call):
        f = node.func
        if isinstance(f, ast.Attribute) and isinstance(f.value, ast.Name) \
                and f.value.id == "np" and f.attr == "array" and len(node.args) == 1:
            return value(node.args[0], env)
        if isinstance(f, ast.Name) and f.id == "Mesh":
            args = [value(a, env) for a in node.args]
            return {"x_intervals": args[0], "d_x": args[1],
                    "coordinate_system": args[2] if len(args) > 2 else "CARTESIAN"}
        raise Skip("call")
    raise Skip(type(node).__name__)


def extract(fn: ast.FunctionDef):
    env = {}
    call = None
    for stmt in fn.body:
        if isinstance(stmt, ast.Assign) and len(stmt.targets) == 1 \
                and isinstance(stmt.targets[0], ast.Name):
            name = stmt.targets[0].id
            rhs = stmt.value
            if isinstance(rhs, ast.Call) and isinstance(rhs.func, ast.Attribute) \
                    and isinstance(rhs.func.value, ast.Name) and rhs.func.value.id == "diff":
                if name.startswith("actual"):
                    call = (rhs.func.attr, rhs)
                continue
            if isinstance(rhs, ast.Call) and isinstance(rhs.func, ast.Name) \
                    and rhs.func.id == "ThreePointCentralDifferenceMethod":
                continue
            env[name] = value(rhs, env)
    if call is None:
        raise Skip("no single operator call")
    op, node = call
    if op not in OPS:
        raise Skip(op)
    args = [value(a, env) for a in node.args[2:]]
    kwargs = {k.arg: value(k.value, env) for k in node.keywords}
    expected = [v for k, v in env.items() if k.startswith("expected")]
    if len(expected) != 1:
        raise Skip("expected array")
    return {"op": op, "mesh": env["mesh"], "y": env["y"], "args": args, "kwargs": kwargs,
            "expected": expected[0]}


def main():
    root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    src = (root / "tests/operators/fdm/test_numerical_differentiator.py").read_text()
    tree = ast.parse(src)
    out = {}
    for fn in tree.body:
        if not isinstance(fn, ast.FunctionDef):
            continue
        m = NAME.match(fn.name)
        if not m or "constrained" in fn.name:
            continue
        try:
            kat = extract(fn)
        except Skip as e:
            print(f"skip {fn.name}: {e}")
            continue
        kat["source"] = f"tests/operators/fdm/test_numerical_differentiator.py:{fn.lineno}"
        out[fn.name] = kat
    dst = Path(__file__).resolve().parent / "curvilinear_kats.json"
    dst.write_text(json.dumps(out, indent=1))
    print(f"wrote {len(out)} cases to {dst}")


if __name__ == "__main__":
    main()
