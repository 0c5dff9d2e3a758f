"""Point-wise evaluation of configuration expressions (the role of
frfs/nputil.py:67-81 npeval with frfs/inifile.py:76-93 getexpr): an
expression of x, y, z, [constants] values and a fixed set of math functions,
evaluated with numpy over arrays of points.  Exponentiation is written
pow(a, b); attribute access and any other character are refused."""
import re

import numpy as np

_SYMS = {
    "sin": np.sin, "cos": np.cos, "tan": np.tan, "asin": np.arcsin, "acos": np.arccos,
    "atan": np.arctan, "atan2": np.arctan2, "sinh": np.sinh, "cosh": np.cosh, "tanh": np.tanh,
    "exp": np.exp, "log": np.log, "log10": np.log10, "sqrt": np.sqrt, "pow": np.power,
    "abs": np.abs, "fabs": np.abs, "min": np.minimum, "max": np.maximum, "pi": np.pi,
}


def npeval(expr, locals):
    if "^" in expr or "**" in expr:
        raise ValueError("Direct exponentiation is not supported; use pow")
    if not re.match(r"[A-Za-z0-9 \t\n\r.,+\-*/%()]+$", expr):
        raise ValueError("Invalid characters in expression")
    names = set(re.findall(r"[A-Za-z_][A-Za-z0-9_]*", re.sub(r"\d+\.?\d*[eE][+-]?\d+", "", expr)))
    unknown = names - set(_SYMS) - set(locals)
    if unknown:
        raise ValueError(f"Unknown names in expression: {sorted(unknown)}")
    if re.search(r"[A-Za-z0-9_)]\s*\.\s*[A-Za-z_]", expr):
        raise ValueError("Invalid expression")
    return eval(expr, {"__builtins__": {}, **_SYMS}, dict(locals))


def eval_at_points(cfg, sect, key, ploc):
    """Value of cfg[sect][key] at points ploc (n, ndims) as an (n,) array."""
    env = dict(cfg.items_as("constants", float))
    for d, name in enumerate("xyz"[: ploc.shape[1]]):
        env[name] = ploc[:, d]
    val = npeval(cfg.get(sect, key), env)
    return np.broadcast_to(np.asarray(val, dtype=float), (ploc.shape[0],)).copy()
