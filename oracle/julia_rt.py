"""Run-time support for the Python translation of the reference's Julia sources.

oracle/ref_julia.py translates the reference's per-cell Julia functions (closure,
realizability, wave speeds, BGK) into oracle/_ref/reference_jl.py; that generated module
does `from julia_rt import *`.  This file supplies the Julia semantics the translation relies
on, written for that purpose (test infrastructure only, never on the product path):

- JArr: an array with Julia's 1-based, column-major indexing (ranges a:b inclusive, linear
  indexing of multi-dimensional arrays, index vectors), `*` as the matrix product;
- NaN-propagating min/max/minimum/maximum, sign, eps, Julia's real/complex sqrt rules
  (DomainError outside @fastmath, NaN inside), det/eigvals/sqrt(::Matrix)/`\\` on LAPACK via
  numpy (the same LAPACK routines Julia calls: getrf, geev, syevr-based sqrt for symmetric
  input);
- NamedTuple / merge, fill, zeros, broadcasting `f.(x)`.
"""
import cmath
import math

import numpy as np

np.seterr(all="ignore")  # IEEE results (Inf, NaN) as Julia gives them, without warnings


class JuliaError(Exception):
    """An exception the Julia code would throw (DomainError, error(...))."""


class _Ref:
    def __init__(self, v):
        self.value = v


POSITIVITY_ENABLED = _Ref(True)
NaN = float("nan")
Inf = float("inf")
pi = math.pi
true = True
false = False
nothing = None
ComplexF64 = complex
Float64 = float


def _fix_key(k, shape, axis):
    if isinstance(k, slice):
        if k.step is not None:
            raise JuliaError("stepped range index not supported")
        if k.start is None and k.stop is None:
            return slice(None)
        return slice(int(k.start) - 1, int(k.stop))
    if isinstance(k, JArr):
        return np.asarray(k.a, dtype=np.int64).ravel(order="F") - 1
    if isinstance(k, (list, tuple, np.ndarray)):
        return np.asarray(k, dtype=np.int64) - 1
    return int(k) - 1


class JArr:
    """Julia array semantics over a numpy array."""

    __array_priority__ = 100
    __array_ufunc__ = None  # numpy scalars defer to the reflected operators below

    def __init__(self, a):
        self.a = np.asarray(a)

    # indexing ---------------------------------------------------------------------------
    def _key(self, key):
        if not isinstance(key, tuple):
            key = (key,)
        if len(key) == 1 and self.a.ndim > 1:  # linear (column-major) indexing
            return None, _fix_key(key[0], None, 0)
        return tuple(_fix_key(k, self.a.shape, i) for i, k in enumerate(key)), None

    def __getitem__(self, key):
        nd, lin = self._key(key)
        if nd is None:
            r = self.a.ravel(order="F")[lin]
        else:
            r = self.a[nd]
        if isinstance(r, np.ndarray) and r.ndim > 0:
            return JArr(r.copy())
        return r  # a numpy scalar: IEEE division by zero, as in Julia

    def __setitem__(self, key, val):
        if isinstance(val, JArr):
            val = val.a
        nd, lin = self._key(key)
        if nd is None:
            flat = self.a.ravel(order="F")
            if isinstance(val, complex) and not np.iscomplexobj(self.a):
                self.a = self.a.astype(complex)
                flat = self.a.ravel(order="F")
            flat[lin] = np.asarray(val).ravel(order="F") if np.ndim(val) else val
            self.a = flat.reshape(self.a.shape, order="F")
        else:
            if (isinstance(val, complex) or np.iscomplexobj(val)) and not np.iscomplexobj(self.a):
                self.a = self.a.astype(complex)
            self.a[nd] = val

    def __len__(self):
        return self.a.size

    def __iter__(self):
        for v in self.a.ravel(order="F"):
            yield v

    # arithmetic -------------------------------------------------------------------------
    @staticmethod
    def _v(x):
        return x.a if isinstance(x, JArr) else x

    def __add__(self, o):
        return JArr(self.a + self._v(o))

    __radd__ = __add__

    def __sub__(self, o):
        return JArr(self.a - self._v(o))

    def __rsub__(self, o):
        return JArr(self._v(o) - self.a)

    def __mul__(self, o):
        if isinstance(o, JArr):
            return _wrap(self.a @ o.a)
        return JArr(self.a * o)

    def __rmul__(self, o):
        if isinstance(o, JArr):
            return _wrap(o.a @ self.a)
        return JArr(o * self.a)

    def __truediv__(self, o):
        if isinstance(o, JArr):
            raise JuliaError("matrix right division not supported")
        return JArr(self.a / o)

    def __neg__(self):
        return JArr(-self.a)

    def __pos__(self):
        return self

    def __eq__(self, o):
        return isinstance(o, JArr) and self.a.shape == o.a.shape and bool(np.all(self.a == o.a))

    def __repr__(self):
        return "JArr(%r)" % (self.a,)


def _wrap(r):
    if isinstance(r, np.ndarray) and r.ndim > 0:
        return JArr(r)
    return r


def _arr(x):
    if isinstance(x, JArr):
        return x.a
    return np.asarray(x)


# constructors ---------------------------------------------------------------------------
def jvec(items):
    """[a, b, c] and [a; b] (vertical concatenation of scalars and vectors)."""
    parts = []
    for v in items:
        if isinstance(v, JArr):
            parts.extend(v.a.ravel(order="F").tolist())
        else:
            parts.append(v)
    return JArr(np.array(parts))


def jmat(rows):
    """[a b; c d]: rows of scalars."""
    return JArr(np.array([[(_arr(v).item() if isinstance(v, JArr) else v) for v in r] for r in rows]))


def jsplat(x):
    return list(x)


def zeros(*dims):
    dt = float
    if dims and dims[0] in (complex, float):
        dt, dims = dims[0], dims[1:]
    if len(dims) == 1 and isinstance(dims[0], tuple):
        dims = dims[0]
    return JArr(np.zeros(tuple(int(d) for d in dims), dtype=dt))


def fill(v, *dims):
    return JArr(np.full(tuple(int(d) for d in dims), v))


def reshape(a, *dims):
    if len(dims) == 1 and isinstance(dims[0], tuple):
        dims = dims[0]
    return JArr(np.reshape(_arr(a) if not isinstance(a, list) else np.array(a), dims, order="F"))


def size(a, d=None):
    s = _arr(a).shape
    return s if d is None else s[d - 1]


def length(a):
    return _arr(a).size


def copy(a):
    return JArr(_arr(a).copy()) if isinstance(a, JArr) else a


def vec(a):
    return JArr(_arr(a).ravel(order="F"))


def jrange(a, b, step=1):
    a, b, step = int(a), int(b), int(step)
    return range(a, b + 1, step) if step > 0 else range(a, b - 1, step)


def div(a, b):
    return int(a) // int(b) if (a >= 0) == (b >= 0) else -(abs(int(a)) // abs(int(b)))


class NT:
    """NamedTuple."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def merge(a, b):
    d = dict(a.__dict__)
    d.update(b.__dict__)
    return NT(**d)


# scalar semantics -----------------------------------------------------------------------
def eps(x=None):
    if x is None:
        return 2.220446049250313e-16
    return float(np.spacing(abs(float(x))))


def _isnan(x):
    return isinstance(x, float) and x != x


def jmax(*a):
    if len(a) == 1:
        return maximum(a[0])
    r = a[0]
    for v in a[1:]:
        if _isnan(r) or _isnan(v):
            return NaN
        if v > r or (v == r == 0 and math.copysign(1, r) < 0):
            r = v
    return r


def jmin(*a):
    if len(a) == 1:
        return minimum(a[0])
    r = a[0]
    for v in a[1:]:
        if _isnan(r) or _isnan(v):
            return NaN
        if v < r or (v == r == 0 and math.copysign(1, v) < 0):
            r = v
    return r


def maximum(x):
    v = _arr(x).ravel()
    return NaN if np.any(np.isnan(v)) else float(np.max(v))


def minimum(x):
    v = _arr(x).ravel()
    return NaN if np.any(np.isnan(v)) else float(np.min(v))


def jsum(x):
    return _wrap(np.sum(_arr(x)))


def sign(x):
    if _isnan(x):
        return NaN
    return (x > 0) - (x < 0) if isinstance(x, int) else float(np.sign(x))


def isfinite(x):
    return bool(np.all(np.isfinite(_arr(x))))


def any_nonfinite(x):
    return not bool(np.all(np.isfinite(_arr(x))))


def jabs(x):
    if isinstance(x, JArr):
        return JArr(np.abs(x.a))
    return abs(x)


def real(x):
    if isinstance(x, JArr):
        return JArr(np.real(x.a))
    return x.real if isinstance(x, complex) else x


def imag(x):
    if isinstance(x, JArr):
        return JArr(np.imag(x.a))
    return x.imag if isinstance(x, complex) else 0.0


def Complex(x):
    return complex(x)


def jsort(x):
    a = _arr(x).ravel(order="F")
    if np.iscomplexobj(a):
        raise JuliaError("sort of complex values")
    return JArr(np.sort(a))


def bcast(f, x):
    """f.(x)"""
    if isinstance(x, JArr):
        flat = [f(v) for v in x]
        return JArr(np.array(flat).reshape(x.a.shape, order="F"))
    return f(x)


def jexp(x):
    if isinstance(x, complex):
        return np.complex128(cmath.exp(x))
    return np.exp(np.float64(x))


def jlog(x):
    if isinstance(x, complex):
        return cmath.log(x)
    if x < 0:
        raise JuliaError("DomainError: log of a negative number")
    return np.float64(math.log(x)) if x > 0 else np.float64(-Inf)


def cbrt(x):
    return np.cbrt(np.float64(x))


def _sqrt_scalar(x, fast):
    if isinstance(x, complex):
        return np.complex128(cmath.sqrt(x))
    if x != x:
        return np.float64(NaN)
    if x < 0:
        if fast:
            return np.float64(NaN)
        raise JuliaError("DomainError: sqrt of a negative number")
    return np.float64(math.sqrt(x))


def _sqrt_matrix(A):
    """Julia's sqrt(::Matrix{Float64}) for a symmetric argument: eigen-decomposition, with
    eigenvalues >= -rtol*max|eig| treated as zeros (real result), complex otherwise."""
    a = A.a
    if not np.array_equal(a, a.T):
        raise JuliaError("sqrt of a nonsymmetric matrix is not translated")
    w, V = np.linalg.eigh(a)
    rtol = np.finfo(float).eps * a.shape[0]
    lam0 = -np.max(np.abs(w)) * rtol
    if np.all(w >= lam0):
        R = (V * np.sqrt(np.maximum(0.0, w))) @ V.T
        # Julia copies the upper triangle into the lower one (copytri!)
        iu = np.triu_indices(a.shape[0], 1)
        R[(iu[1], iu[0])] = R[iu]
        return JArr(R)
    R = (V * np.sqrt(w.astype(complex))) @ V.T
    return JArr(R)


def jsqrt(x):
    if isinstance(x, JArr):
        return _sqrt_matrix(x)
    return _sqrt_scalar(x, False)


def jsqrt_fast(x):
    if isinstance(x, JArr):
        return _sqrt_matrix(x)
    return _sqrt_scalar(x, True)


def jpow(a, b):
    """a^b with Julia's rules: a negative real base with a non-integer exponent throws."""
    if isinstance(a, JArr):
        if isinstance(b, int) and a.a.ndim == 2:
            return JArr(np.linalg.matrix_power(a.a, b))
        raise JuliaError("array power")
    if isinstance(a, complex) or isinstance(b, complex):
        return complex(a) ** b
    if isinstance(b, int) or (isinstance(b, float) and b.is_integer() and abs(b) < 2**53):
        return a ** b if isinstance(b, int) else float(a) ** b
    if a < 0:
        raise JuliaError("DomainError: negative base with a fractional exponent")
    return float(a) ** b


# linear algebra -------------------------------------------------------------------------
def det(A):
    return float(np.linalg.det(A.a))


def dot(a, b):
    return _wrap(np.vdot(_arr(a).ravel(order="F"), _arr(b).ravel(order="F")))


def ldiv(A, b):
    return JArr(np.linalg.solve(A.a, _arr(b)))


def eigvals(A):
    if not np.all(np.isfinite(A.a)):
        raise JuliaError("ArgumentError: matrix contains Infs or NaNs")
    return JArr(np.linalg.eigvals(A.a))


def jprint(*a, **k):
    pass
