"""Recipe: translate the reference's per-cell Julia functions into Python, as a checker for the
oracle.

The reference (comp-physics/HyQMOM.jl) is Julia, and no Julia runtime is available here.
This script translates the subset of Julia that the per-cell numerics use (assignments,
if/elseif/else, for/while, tuple returns and destructuring, array literals, 1-based
indexing, a few macros) into Python, statement by statement, and writes
oracle/_ref/reference_jl.py (git-ignored; never committed and never shipped).  Array and
scalar semantics (1-based column-major indexing, NaN-propagating min/max, DomainError on
sqrt of a negative outside @fastmath, LAPACK det/eigvals/sqrt) come from oracle/julia_rt.py.

The translation is strict: a construct outside the subset raises at translation time rather
than being guessed.  Only test-vector scripts under tests/golden/ import the output, in this
container; the GPU tests and the product never do.

Usage:  python oracle/ref_julia.py [/root/reference]
"""
import os
import re
import sys

# (file under src/, functions to take; None = every function in the file)
SOURCES = [
    ("autogen/M4toC4_3D.jl", None),
    ("autogen/C4toM4_3D.jl", None),
    ("autogen/C5toM5_3D.jl", None),
    ("autogen/S4toC4_3D_r.jl", None),
    ("autogen/delta2star3D.jl", None),
    ("autogen/delta2star3D_permutation.jl", None),
    ("autogen/lower_bound_S220.jl", None),
    ("autogen/bound_minor1.jl", None),
    ("autogen/delta2starchol_L3.jl", None),
    ("autogen/rootsR.jl", None),
    ("autogen/rootsR_X_Y.jl", None),
    ("autogen/jacobian6.jl", None),
    ("moments/M2CS4_35.jl", None),
    ("moments/enforce_univariate.jl", None),
    ("moments/hyqmom_3D.jl", None),
    ("moments/Moments5_3D.jl", None),
    ("moments/InitializeM4_35.jl", None),
    ("realizability/realizability.jl", ["realizability"]),
    ("realizability/realizability_S111.jl", None),
    ("realizability/realizability_S2.jl", None),
    ("realizability/realizability_S210.jl", None),
    ("realizability/realizability_S211.jl", None),
    ("realizability/realizability_S310.jl", None),
    ("realizability/realizability_S310_220.jl", None),
    ("realizability/realizablity_S220.jl", None),
    ("realizability/realizable_2D.jl", None),
    ("realizability/realizable_3D.jl", None),
    ("realizability/edge_corner_correction.jl", None),
    ("utils/diagnostics.jl", ["check2D", "check2D_all_planes"]),
    ("utils/moment_array_utils.jl", ["M4_to_vars", "M5_to_vars", "S_to_C_batch"]),
    ("utils/axis_moment_slice.jl", None),
    ("numerics/Flux_closure35_and_realizable_3D.jl", None),
    ("numerics/compute_jacobian_eigenvalues.jl", None),
    ("numerics/eigenvalues6_hyperbolic_3D.jl", None),
    ("numerics/eigenvalues6z_hyperbolic_3D.jl", None),
    ("numerics/closure_and_eigenvalues.jl", None),
    ("numerics/collision35.jl", None),
    ("numerics/flux_HLL.jl", None),
    ("numerics/pas_HLL.jl", None),
]

FUNC_MAP = {  # Julia function -> julia_rt name (called with parentheses)
    "max": "jmax", "min": "jmin", "exp": "jexp", "log": "jlog", "abs": "jabs",
    "sort": "jsort", "sum": "jsum", "println": "jprint", "print": "jprint",
}
OPEN_KW = ("function", "if", "for", "while", "begin", "let")


class TranslateError(Exception):
    pass


# ---------------------------------------------------------------------------------------
# lexical helpers
def _split_strings(s):
    """Split s into [(is_string, text)] segments (Julia double-quoted strings)."""
    out, i, buf = [], 0, ""
    while i < len(s):
        if s[i] == '"':
            if buf:
                out.append((False, buf))
                buf = ""
            j = i + 1
            while j < len(s) and s[j] != '"':
                j += 2 if s[j] == "\\" else 1
            out.append((True, s[i:j + 1]))
            i = j + 1
        else:
            buf += s[i]
            i += 1
    if buf:
        out.append((False, buf))
    return out


def _strip_comment(line):
    res = ""
    for is_str, t in _split_strings(line):
        if is_str:
            res += t
        else:
            k = t.find("#")
            if k >= 0:
                return res + t[:k]
            res += t
    return res


def _depth_delta(line):
    d = 0
    for is_str, t in _split_strings(line):
        if not is_str:
            d += sum(t.count(c) for c in "([{") - sum(t.count(c) for c in ")]}")
    return d


def logical_lines(src):
    """Comments and docstrings removed; continuation lines joined."""
    lines = src.split("\n")
    out, i, in_doc = [], 0, False
    raw = []
    for ln in lines:
        s = ln.strip()
        if s.startswith('"""'):
            if not (len(s) > 3 and s.endswith('"""')):
                in_doc = not in_doc
            continue
        if in_doc:
            continue
        raw.append(_strip_comment(ln).rstrip())
    buf, depth = "", 0
    for ln in raw:
        if not ln.strip():
            if buf and depth == 0:
                out.append(buf)
                buf = ""
            continue
        buf = (buf + " " + ln.strip()) if buf else ln.strip()
        depth += _depth_delta(ln)
        if depth > 0 or re.search(r"(,|=|\+|-|\*|/|&&|\|\|)$", buf):
            continue
        out.append(buf)
        buf = ""
    if buf:
        out.append(buf)
    return out


def split_statements(line):
    """Split on ';' at bracket depth 0 (a ';' in (...) separates keyword arguments and is
    turned into ','; inside [...] it is a row separator and is kept)."""
    parts, buf, stack = [], "", []
    for is_str, t in _split_strings(line):
        if is_str:
            buf += t
            continue
        for ch in t:
            if ch in "([{":
                stack.append(ch)
            elif ch in ")]}":
                stack.pop()
            if ch == ";" and not stack:
                parts.append(buf.strip())
                buf = ""
                continue
            if ch == ";" and stack and stack[-1] == "(":
                buf += ","
                continue
            buf += ch
    if buf.strip():
        parts.append(buf.strip())
    return [p for p in parts if p]


# ---------------------------------------------------------------------------------------
# expressions
def _find_close(s, i):
    """s[i] is an opening bracket; index of the matching closer."""
    d = 0
    for j in range(i, len(s)):
        if s[j] in "([{":
            d += 1
        elif s[j] in ")]}":
            d -= 1
            if d == 0:
                return j
    raise TranslateError("unbalanced brackets in %r" % s)


def _split_top(s, sep):
    out, d, buf = [], 0, ""
    for ch in s:
        if ch in "([{":
            d += 1
        elif ch in ")]}":
            d -= 1
        if ch == sep and d == 0:
            out.append(buf)
            buf = ""
        else:
            buf += ch
    out.append(buf)
    return out


def _split_ws_elements(row):
    """Elements of a space-separated matrix row, Julia-style: 'a-b  c*d' -> [a-b, c*d];
    an operator with spaces on both sides stays binary."""
    chunks, d, buf = [], 0, ""
    for ch in row.strip():
        if ch in "([{":
            d += 1
        elif ch in ")]}":
            d -= 1
        if ch.isspace() and d == 0:
            if buf:
                chunks.append(buf)
                buf = ""
        else:
            buf += ch
    if buf:
        chunks.append(buf)
    out = []
    for c in chunks:
        if out and (re.search(r"[+\-*/^=<>]$", out[-1]) or re.fullmatch(r"[+\-*/^]", c)
                    or (re.match(r"[*/^]", c))):
            out[-1] = out[-1] + " " + c
        else:
            out.append(c)
    return out


def _array_literal(inner):
    inner = inner.strip()
    if not inner:
        return "jvec([])"
    rows = _split_top(inner, ";")
    if len(rows) == 1 and len(_split_top(inner, ",")) > 1:
        return "jvec([%s])" % ", ".join(translate_expr(e) for e in _split_top(inner, ","))
    elems = [_split_ws_elements(r) for r in rows if r.strip()]
    if all(len(r) == 1 for r in elems):
        return "jvec([%s])" % ", ".join(translate_expr(r[0]) for r in elems)
    if len(set(len(r) for r in elems)) != 1:
        raise TranslateError("ragged matrix literal: %r" % inner)
    return "jmat([%s])" % ", ".join("[%s]" % ", ".join(translate_expr(e) for e in r) for r in elems)


def _rewrite_brackets(s):
    """Array literals -> jvec/jmat; indexing kept (recursing into the index)."""
    out, i = "", 0
    while i < len(s):
        ch = s[i]
        if ch == "[":
            j = _find_close(s, i)
            prev = out.rstrip()
            is_index = bool(prev) and (prev[-1].isalnum() or prev[-1] in "_)]") and \
                not re.search(r"\b(return|in|and|or|not|if|elif|else)$", prev)
            inner = s[i + 1:j]
            if is_index:
                out += "[" + _index_expr(inner) + "]"
            else:
                out += _array_literal(inner)
            i = j + 1
        elif ch == "(":
            j = _find_close(s, i)
            out += "(" + _rewrite_brackets(s[i + 1:j]) + ")"
            i = j + 1
        else:
            out += ch
            i += 1
    return out


def _index_expr(inner):
    parts = _split_top(inner, ",")
    res = []
    for p in parts:
        p = p.strip()
        if p == ":":
            res.append(":")
            continue
        rng = _split_top(p, ":")
        if len(rng) == 2:
            res.append("%s:%s" % (translate_expr(rng[0]), translate_expr(rng[1])))
        elif len(rng) == 1:
            res.append(translate_expr(p))
        else:
            raise TranslateError("stepped index range %r" % p)
    return ", ".join(res)


_FAST = [False]


def _symbols(t):
    """:name -> "name" where the ':' does not continue a range (a:b)."""
    out, i = "", 0
    for m in re.finditer(r":([A-Za-z_]\w*)", t):
        k = m.start() - 1
        while k >= 0 and t[k] == " ":
            k -= 1
        if k >= 0 and (t[k].isalnum() or t[k] in "_)]:"):
            continue
        out += t[i:m.start()] + '"%s"' % m.group(1)
        i = m.end()
    return out + t[i:]


def translate_expr(e):
    e = e.strip()
    if not e:
        return e
    segs = []
    for is_str, t in _split_strings(e):
        if is_str:
            segs.append(t.replace("$", ""))
            continue
        t = re.sub(r"(\w+)\.\.\.", r"*\1", t)
        t = re.sub(r"(?<![\w.])(\d+(?:\.\d+)?)im\b", r"\1j", t)
        t = re.sub(r"\bim\b", "1j", t)
        t = t.replace("POSITIVITY_ENABLED[]", "POSITIVITY_ENABLED.value")
        t = re.sub(r"\bany\(\s*!\s*isfinite\s*,", "any_nonfinite(", t)
        t = _symbols(t)
        if "'" in t:
            raise TranslateError("transpose/char literal in %r" % e)
        if "?" in t:
            raise TranslateError("ternary in %r" % e)
        for m in re.finditer(r"\.?\^\s*(\(?[^\s*/+\-),\]]*\)?)", t):
            if not re.fullmatch(r"\(?-?\d+\)?", m.group(1)):
                raise TranslateError("non-integer power %r in %r" % (m.group(1), e))
        t = re.sub(r"\.\^", "**", t).replace("^", "**")
        t = re.sub(r"(\w+)\s*\\\s*(\w+)", r"ldiv(\1, \2)", t)
        if "\\" in t:
            raise TranslateError("left division in %r" % e)
        t = re.sub(r"\.([+\-*/])", r"\1", t)  # .+ .- .* ./ on arrays of equal shape
        t = t.replace("&&", " and ").replace("||", " or ")
        t = re.sub(r"!(?!=)", " not ", t)

        def fmap(m):
            name = m.group(1)
            if name == "sqrt":
                return ("jsqrt_fast" if _FAST[0] else "jsqrt") + m.group(2)
            return FUNC_MAP.get(name, name) + m.group(2)

        t = re.sub(r"(?<![\w.])(\w+)(\.?\()", fmap, t)
        t = re.sub(r"(?<![\w.])(\w+)\.\(", r"bcast(\1, ", t)
        segs.append(t)
    e = "".join(segs)
    e = re.sub(r'\bSymbol\(\s*("[^"]*")\s*\)', r"\1", e)
    e = re.sub(r"(?<![\w\])])\(\s*(\w+)\s*=(?!=)", r"NT(\1=", e)
    return _rewrite_brackets(e)


# ---------------------------------------------------------------------------------------
# statements
def _strip_types(params):
    out = []
    for p in _split_top(params, ","):
        p = p.strip()
        if not p:
            continue
        p = re.sub(r"::[\w.]+(\{[^}]*\})?", "", p)
        if p.endswith("..."):
            p = "*" + p[:-3]
        out.append(p)
    return out


def _first_param_ndim(params):
    first = _split_top(params.split(";")[0], ",")[0]
    if "AbstractArray{T,3}" in first or "Array{Float64,3}" in first:
        return 3
    if "AbstractVector" in first or "Vector" in first:
        return 1
    return None


def _tuple_return(expr):
    x = expr.strip()
    if re.match(r"^\(\s*\w+\s*=(?!=)", x):  # NamedTuple
        return translate_expr(x)
    if x.startswith("(") and _find_close(x, 0) == len(x) - 1 and len(_split_top(x[1:-1], ",")) > 1:
        return "JTuple((%s))" % translate_expr(x[1:-1])
    if len(_split_top(x, ",")) > 1:
        return "JTuple((%s))" % translate_expr(x)
    return translate_expr(x)


def translate_function(stmts, name_override=None):
    """stmts: statements from 'function ...' to its matching 'end'."""
    py, ind, stack = [], 0, []
    first = True
    body_last = None
    for idx, st in enumerate(stmts):
        fast_here = False
        if st.startswith("@fastmath "):
            st = st[len("@fastmath "):].strip()
            fast_here = True
        if st.startswith("@inbounds "):
            st = st[len("@inbounds "):].strip()
        kw = st.split()[0] if st.split() else ""
        kw = re.match(r"^[\w@]+", st).group(0) if re.match(r"^[\w@]+", st) else kw
        if first:
            m = re.match(r"function\s+(\w+)\s*\((.*)\)\s*(where.*)?$", st)
            if not m:
                raise TranslateError("bad function header %r" % st)
            name = name_override or m.group(1)
            params = m.group(2)
            pos, _, kws = params.partition(";")
            args = _strip_types(pos) + ([translate_expr(k) for k in _strip_types(kws)] if kws else [])
            py.append("def %s(%s):" % (name, ", ".join(args)))
            stack.append(("function", fast_here))
            _FAST[0] = fast_here
            ind = 1
            py.append("    pass")
            first = False
            continue
        pad = "    " * ind
        if kw == "end":
            k, f = stack.pop()
            ind -= 1
            _FAST[0] = any(x[1] for x in stack)
            if not stack:
                break
            continue
        if kw in ("if", "while"):
            py.append(pad + "%s %s:" % (kw, translate_expr(st[len(kw):])))
            stack.append((kw, fast_here or _FAST[0]))
            _FAST[0] = stack[-1][1]
            ind += 1
            py.append("    " * ind + "pass")
            continue
        if kw == "elseif":
            py.append("    " * (ind - 1) + "elif %s:" % translate_expr(st[6:]))
            py.append(pad + "pass")
            continue
        if kw == "else":
            py.append("    " * (ind - 1) + "else:")
            py.append(pad + "pass")
            continue
        if kw == "begin":
            py.append(pad + "if True:")
            stack.append(("begin", fast_here or _FAST[0]))
            _FAST[0] = stack[-1][1]
            ind += 1
            py.append("    " * ind + "pass")
            continue
        if kw == "for":
            m = re.match(r"for\s+(.+?)\s*(?:=|\bin\b)\s*(.+)$", st)
            if not m:
                raise TranslateError("bad for %r" % st)
            var, rng = m.group(1), m.group(2).strip()
            parts = _split_top(rng, ":")
            if len(parts) == 2:
                it = "jrange(%s, %s)" % (translate_expr(parts[0]), translate_expr(parts[1]))
            elif len(parts) == 3:
                it = "jrange(%s, %s, %s)" % (translate_expr(parts[0]), translate_expr(parts[2]),
                                             translate_expr(parts[1]))
            else:
                it = translate_expr(rng)
            py.append(pad + "for %s in %s:" % (var, it))
            stack.append(("for", fast_here or _FAST[0]))
            ind += 1
            py.append("    " * ind + "pass")
            continue
        if kw in ("@warn", "@info", "@printf", "@debug", "@assert", "@show"):
            py.append(pad + "pass")
            continue
        if kw in ("local", "global"):
            st = st[len(kw):].strip()
        if kw == "return":
            py.append(pad + "return " + _tuple_return(st[6:]) if st[6:].strip() else pad + "return None")
            continue
        if kw == "error":
            py.append(pad + "raise JuliaError(%s)" % translate_expr(st[5:]))
            continue
        if kw in ("function", "let", "try", "struct", "module", "do", "macro", "quote"):
            raise TranslateError("unsupported block %r" % st)
        m = re.match(r"^(\w+)\(([\w, ]*)\)\s*=(?!=)\s*(.+)$", st)
        if m:  # local one-line function
            py.append(pad + "%s = lambda %s: (%s)" % (m.group(1), m.group(2), translate_expr(m.group(3))))
            continue
        # assignment / expression
        m = re.match(r"^(.+?)\s*(\.=|\+=|-=|\*=|/=|(?<![=!<>.])=(?!=))\s*(.+)$", st)
        if m and _depth_delta(m.group(1)) == 0:
            lhs = m.group(1).strip()
            if m.group(2) == ".=":  # broadcast assignment into an existing array
                lhs = lhs if lhs.endswith("]") else lhs + "[:]"
                py.append(pad + "%s = %s" % (translate_expr(lhs), translate_expr(m.group(3))))
                continue
            if lhs.startswith("(") and _find_close(lhs, 0) == len(lhs) - 1:
                lhs = lhs[1:-1]
            lhs_py = ", ".join(translate_expr(x) if x.strip() != "_" else "_" for x in _split_top(lhs, ","))
            py.append(pad + "%s %s %s" % (lhs_py, m.group(2), translate_expr(m.group(3))))
        else:
            # the value of the last expression of a function is its return value
            nxt = stmts[idx + 1] if idx + 1 < len(stmts) else "end"
            if nxt == "end" and len(stack) == 1:
                py.append(pad + "return " + _tuple_return(st))
            else:
                py.append(pad + translate_expr(st))
    if stack:
        raise TranslateError("unterminated block in function")
    return "\n".join(py) + "\n"


def functions_in(src):
    """[(name, first-parameter ndim hint, statements)] of each top-level function."""
    stmts = []
    for ln in logical_lines(src):
        stmts.extend(split_statements(ln))
    out, i = [], 0
    while i < len(stmts):
        st = stmts[i]
        s2 = re.sub(r"^@fastmath\s+", "", st)
        if s2.startswith("function "):
            depth, j = 0, i
            while j < len(stmts):
                t = re.sub(r"^@(fastmath|inbounds)\s+", "", stmts[j])
                w = re.match(r"^\w+", t)
                w = w.group(0) if w else ""
                if w in OPEN_KW:
                    depth += 1
                elif w == "end":
                    depth -= 1
                    if depth == 0:
                        break
                j += 1
            m = re.match(r"function\s+(\w+)\s*\((.*)\)", s2)
            out.append((m.group(1), _first_param_ndim(m.group(2)), stmts[i:j + 1]))
            i = j + 1
        else:
            i += 1
    return out


PRELUDE = '''# Generated by oracle/ref_julia.py from the reference's Julia sources; do not commit.
import os, sys
sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from julia_rt import *  # noqa: F401,F403
from julia_rt import JuliaError


class JTuple(tuple):
    """A Julia tuple: 1-based indexing."""
    def __getitem__(self, k):
        return tuple.__getitem__(self, k - 1)

'''


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    outdir = os.path.join(here, "_ref")
    os.makedirs(outdir, exist_ok=True)
    parts = [PRELUDE]
    methods = {}
    for rel, names in SOURCES:
        with open(os.path.join(ref, "src", rel)) as f:
            src = f.read()
        for name, nd, stmts in functions_in(src):
            if names is not None and name not in names:
                continue
            methods.setdefault(name, []).append((rel, nd, stmts))
    for name, ms in methods.items():
        if len(ms) == 1:
            rel, _, stmts = ms[0]
            parts.append("# src/%s\n" % rel + translate_function(stmts))
            continue
        disp = ["def %s(*args, **kw):" % name]
        for k, (rel, nd, stmts) in enumerate(ms):
            mname = "%s__m%d" % (name, k)
            parts.append("# src/%s\n" % rel + translate_function(stmts, mname))
            if nd is None:
                raise TranslateError("cannot dispatch %s" % name)
            disp.append("    if isinstance(args[0], JArr) and args[0].a.ndim == %d:" % nd)
            disp.append("        return %s(*args, **kw)" % mname)
        disp.append("    raise JuliaError('no method of %s')" % name)
        parts.append("\n".join(disp) + "\n")
    path = os.path.join(outdir, "reference_jl.py")
    with open(path, "w") as f:
        f.write("\n\n".join(parts))
    print("wrote", path)


if __name__ == "__main__":
    main()
