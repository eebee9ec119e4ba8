"""The oracle's math-derived forms of the reference's symbolic-toolbox functions against
vectors evaluated with the reference formulas (tests/golden/autogen_vectors.json)."""
import numpy as np
import pytest

import oracle as O
from _common import E35, load_json

V = load_json("autogen_vectors.json")
RTOL = 1e-11


def _close(a, b, rtol=RTOL, atol_scale=1e-12):
    a = np.asarray(a, float); b = np.asarray(b, float)
    scale = max(1.0, np.abs(b).max())
    assert np.all(np.abs(a - b) <= rtol * np.abs(b) + atol_scale * scale), (a, b)


@pytest.mark.parametrize("i", range(12))
def test_m4toc4(i):
    v = V["M4toC4_3D"][i]
    C, _ = O.m2cs4(v["in"])
    ref = np.array(v["out"]).reshape((5, 5, 5), order="F")
    _close(C, [ref[e] for e in E35])


@pytest.mark.parametrize("i", range(12))
def test_c5tom5(i):
    v = V["C5toM5_3D"][i]
    a = v["in"]
    M000, u, vv, w = a[:4]
    names4 = [(2,0,0),(1,1,0),(1,0,1),(0,2,0),(0,1,1),(0,0,2),(3,0,0),(2,1,0),(2,0,1),(1,2,0),
              (1,1,1),(1,0,2),(0,3,0),(0,2,1),(0,1,2),(0,0,3),(4,0,0),(3,1,0),(3,0,1),(2,2,0),
              (2,1,1),(2,0,2),(1,3,0),(1,2,1),(1,1,2),(1,0,3),(0,4,0),(0,3,1),(0,2,2),(0,1,3),(0,0,4)]
    names5 = [(5,0,0),(4,1,0),(3,2,0),(2,3,0),(1,4,0),(4,0,1),(3,0,2),(2,0,3),(1,0,4),(3,1,1),
              (2,2,1),(1,3,1),(2,1,2),(1,1,3),(1,2,2),(0,5,0),(0,4,1),(0,3,2),(0,2,3),(0,1,4),(0,0,5)]
    C = np.zeros((6, 6, 6))
    for e, x in zip(names4 + names5, a[4:]):
        C[e] = x
    C[0, 0, 0] = 1.0
    M = np.zeros(216)
    O.lib().orc_central_to_raw5(O.ctypes.c_double(M000), O.ctypes.c_double(u), O.ctypes.c_double(vv),
                                O.ctypes.c_double(w), O._p(np.ascontiguousarray(C.ravel())), O._p(M))
    M = M.reshape(6, 6, 6)
    ref = np.array(v["out"]).reshape((6, 6, 6), order="F")
    for i in range(6):
        for j in range(6 - i):
            for k in range(6 - i - j):
                _close(M[i, j, k], ref[i, j, k])


@pytest.mark.parametrize("i", range(12))
def test_delta2star3D(i):
    v = V["delta2star3D"][i]
    _close(O.delta2star3D(v["in"]).ravel(), v["out"], rtol=1e-10)


@pytest.mark.parametrize("i", range(12))
def test_bounds(i):
    v = V["lower_bound_S220"][i]
    _close(O.lower_bound_S220(v["in"]), v["out"], rtol=1e-10)
    v = V["bound_minor1"][i]
    _close(O.bound_minor1(v["in"]), v["out"], rtol=1e-10)


@pytest.mark.parametrize("i", range(12))
def test_L3(i):
    v = V["delta2starchol_L3"][i]
    _close(O.delta2starchol_L3(v["in"]), v["out"], rtol=1e-9)


@pytest.mark.parametrize("i", range(12))
def test_cardano(i):
    v = V["rootsR_X_Y"][i]
    R = O.rootsR_X_Y(v["in"])
    ref = np.array(v["re"]) + 1j * np.array(v["im"])
    np.testing.assert_allclose(np.sort_complex(R), np.sort_complex(ref), rtol=1e-8, atol=1e-9)
    v = V["rootsR"][i]
    R = O.rootsR(v["in"])
    ref = np.array(v["re"]) + 1j * np.array(v["im"])
    np.testing.assert_allclose(np.sort_complex(R), np.sort_complex(ref), rtol=1e-8, atol=1e-9)


def test_cardano_three_real_roots_is_nan():
    for e in V["rootsR_X_Y_nan_inputs"]:
        assert np.any(np.isnan(O.rootsR_X_Y(e)))


@pytest.mark.parametrize("i", range(36))
def test_jacobian6(i):
    v = V["jacobian6"][i]
    J = O.jacobian6(v["in"])
    ref = np.array(v["out"]).reshape(6, 6)
    np.testing.assert_allclose(J, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())
