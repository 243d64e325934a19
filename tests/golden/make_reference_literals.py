"""Extract the literal expected arrays of the reference's own tests into a
data fixture (reference_literals.npz).

Run once in a container that holds the reference checkout (path given as the
first argument, default /root/reference).  The test files are parsed as text
with ``ast`` (nothing from the reference is imported or executed); only the
numeric literals are kept:

  tests/test_flows.py:68-163   prefilter coefficients (get_interp_arrays_2D,
                               get_interp_arrays_2D_steady)
  tests/test_flows.py:208-258  tricubic evaluation at 27 in-grid points
  tests/test_flows.py:261-285  trilinear evaluation at the same 27 points
  tests/test_flows.py:307-357  double-gyre / Bickley velocities at 27 points
  tests/test_utils.py          gradF_stencil_2D and eigvalsh_max_2D values
"""
import ast
import os
import sys

import numpy as np

ROOT = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_literals.npz")


def _literal(node):
    """Evaluate np.array([...]) / tuples of them / plain literals."""
    if isinstance(node, ast.Call) and getattr(node.func, "attr", None) == "array":
        return np.array(ast.literal_eval(node.args[0]), dtype=np.float64)
    if isinstance(node, ast.Tuple):
        vals = [_literal(e) for e in node.elts]
        if all(isinstance(v, np.ndarray) for v in vals):
            return np.stack(vals)
        return np.array(vals, dtype=np.float64)
    return np.array(ast.literal_eval(node), dtype=np.float64)


def _assignments(path, func_name, names):
    tree = ast.parse(open(path).read())
    for fn in ast.walk(tree):
        if isinstance(fn, ast.FunctionDef) and fn.name == func_name:
            out = {}
            for st in fn.body:
                if isinstance(st, ast.Assign) and isinstance(st.targets[0], ast.Name):
                    name = st.targets[0].id
                    if name in names:
                        out[name] = _literal(st.value)
            return out
    raise KeyError(func_name)


def main():
    tf = os.path.join(ROOT, "tests", "test_flows.py")
    tu = os.path.join(ROOT, "tests", "test_utils.py")
    d = {}
    a = _assignments(tf, "test_get_interp_arrays_2D", ("Cu_expected", "Cv_expected"))
    d["prefilter3_Cu"] = a["Cu_expected"]
    d["prefilter3_Cv"] = a["Cv_expected"]
    a = _assignments(tf, "test_get_interp_arrays_2D_steady", ("Cu_expected", "Cv_expected"))
    d["prefilter2_Cu"] = a["Cu_expected"]
    d["prefilter2_Cv"] = a["Cv_expected"]
    a = _assignments(tf, "test_get_callable_2D", ("veli_expected",))
    d["spline3_vel"] = a["veli_expected"]
    a = _assignments(tf, "test_get_callable_linear_2D", ("veli_expected",))
    d["linear3_vel"] = a["veli_expected"]
    a = _assignments(tf, "test_get_predefined_callable_dg", ("veli_expected", "domain_expected"))
    d["dg_vel"] = a["veli_expected"]
    d["dg_domain"] = a["domain_expected"]
    a = _assignments(tf, "test_get_predefined_callable_bickley", ("veli_expected",))
    d["bickley_vel"] = a["veli_expected"]
    a = _assignments(tu, "test_gradF_stencil_2D", ("gradF_expected",))
    d["gradF_5_5"] = a["gradF_expected"]
    a = _assignments(tu, "test_eigvalsh_max_2D", ("eigval_expected",))
    d["eigvalsh_max"] = a["eigval_expected"]
    np.savez(OUT, **d)
    for k, v in d.items():
        print(k, v.shape)


if __name__ == "__main__":
    main()
