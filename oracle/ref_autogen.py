"""Recipe: build oracle/_ref/autogen_ref.py from the reference's own symbolic-toolbox
sources (src/autogen/*.jl of comp-physics/HyQMOM.jl).

The autogen files are straight-line FP64 arithmetic.  This script translates that subset
of Julia syntax to Python so the reference formulas themselves can be evaluated as a
checker for the oracle (the analogue of compiling a C reference from its sources).  The
output goes to oracle/_ref/ (git-ignored); nothing of the reference is committed.

Usage:  python oracle/ref_autogen.py [/root/reference]
"""
import os
import re
import sys

FUNCS = [
    ("M4toC4_3D.jl", "M4toC4_3D"),
    ("C4toM4_3D.jl", "C4toM4_3D"),
    ("C5toM5_3D.jl", "C5toM5_3D"),
    ("delta2star3D.jl", "delta2star3D"),
    ("lower_bound_S220.jl", "lower_bound_S220"),
    ("bound_minor1.jl", "bound_minor1"),
    ("delta2starchol_L3.jl", "delta2starchol_L3"),
    ("rootsR.jl", "rootsR"),
    ("rootsR_X_Y.jl", "rootsR_X_Y"),
    ("jacobian6.jl", "jacobian6"),
]


def translate(src, fname):
    m = re.search(r"function\s+%s\((.*?)\)(.*?)\nend" % fname, src, re.S)
    if m is None:
        raise ValueError("function %s not found" % fname)
    args = [a.strip() for a in m.group(1).replace("\n", " ").split(",")]
    lines = []
    for ln in m.group(2).split("\n"):
        ln = ln.split("#")[0].rstrip().strip()
        if not ln or ln in ("@fastmath begin", "begin", "end"):
            continue
        ln = ln.replace("^", "**").replace("0.5im", "0.5j")
        ln = re.sub(r"(\w+)\.\.\.", r"*\1", ln)
        if ln.startswith("return"):
            lines.append(ln)
            continue
        for part in ln.split(";"):
            if part.strip():
                lines.append(part.strip())
    out, acc = [], None
    for l in lines:  # multi-line matrix literal  E = [a b ...; ...]
        if acc is None and re.match(r"^\w+ = \[\w+ \w+", l):
            acc = [l.split("=")[0].strip(), l.split("[")[1].split()]
            continue
        if acc is not None:
            done = "]" in l
            acc[1] += l.replace("]", "").split()
            if done:
                n = int(len(acc[1]) ** 0.5)
                rows = ",".join("[" + ",".join(acc[1][r * n:(r + 1) * n]) + "]" for r in range(n))
                out.append("%s = np.array([%s])" % (acc[0], rows))
                acc = None
            continue
        out.append(l)
    body = "\n".join("    " + l for l in out)
    return ("def %s(%s):\n    from numpy import sqrt, cbrt\n%s\n" % (fname, ", ".join(args), body))


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    outdir = os.path.join(here, "_ref")
    os.makedirs(outdir, exist_ok=True)
    parts = ["# generated from the reference's src/autogen/*.jl; do not commit\n",
             "import numpy as np\n",
             "def reshape(a, *s):\n    return np.reshape(np.array(a), s, order='F')\n\n"]
    for fn, name in FUNCS:
        with open(os.path.join(ref, "src", "autogen", fn)) as f:
            parts.append(translate(f.read(), name))
    with open(os.path.join(outdir, "autogen_ref.py"), "w") as f:
        f.write("\n".join(parts))
    print("wrote", os.path.join(outdir, "autogen_ref.py"))


if __name__ == "__main__":
    main()
