"""The oracle against the reference's MATLAB golden files (tests/golden/mat_goldens.json)
and its own known-answer cases (test/test_golden_files.jl:18-304 uses atol 1e-10)."""
import numpy as np
import pytest

import oracle as O
from _common import load_json

G = load_json("mat_goldens.json")
ATOL = 1e-10


def test_moment_conversions_gaussian():
    g = G["test_moment_conversions_golden"]["gaussian"]
    i = g["input"]
    M = O.initialize_m4_35(i["rho"], i["u"], i["v"], i["w"], i["T"], 0, 0, i["T"], 0, i["T"])
    np.testing.assert_allclose(M, g["output"]["M"], atol=ATOL)
    C, S = O.m2cs4(M)
    np.testing.assert_allclose(C, g["output"]["C4"], atol=ATOL)
    np.testing.assert_allclose(S, g["output"]["S4"], atol=ATOL)


def test_moment_conversions_correlated():
    g = G["test_moment_conversions_golden"]["correlated"]
    i = g["input"]
    M = O.initialize_m4_35(i["rho"], i["u"], i["v"], i["w"], i["C200"], i["C110"], i["C101"],
                           i["C020"], i["C011"], i["C002"])
    np.testing.assert_allclose(M, g["output"]["M"], atol=ATOL)
    C, S = O.m2cs4(M)
    np.testing.assert_allclose(C, g["output"]["C4"], atol=ATOL)
    np.testing.assert_allclose(S, g["output"]["S4"], atol=ATOL)


@pytest.mark.parametrize("case", ["case1", "case2", "case3"])
def test_initialization(case):
    g = G["test_initialization_golden"][case]
    i = g["input"]
    T = i["T"]
    r = np.atleast_1d(i["corr"]).astype(float)
    C110, C101, C011 = r[0] * T, r[1] * T, r[2] * T
    M = O.initialize_m4_35(i["rho"], i["u"], i["v"], i["w"], T, C110, C101, T, C011, T)
    np.testing.assert_allclose(M, g["output"], atol=ATOL)


def test_closure_1d():
    g = G["test_closures_golden"]["closure_1d"]
    Mp, vmin, vmax = O.closure_and_eigenvalues(g["input"])
    assert abs(Mp - g["output"]["Mp"]) < ATOL
    assert abs(vmin - g["output"]["vpmin"]) < ATOL
    assert abs(vmax - g["output"]["vpmax"]) < ATOL


def test_flux_closure_basic():
    g = G["test_flux_eigenvalues_golden"]["flux_basic"]
    # test/test_golden_files.jl:207-221: flag2D=1, Ma=0.5
    Fx, Fy, Fz, Mr, _ = O.flux_closure35(g["input"], 1, 0.5)
    o = g["output"]
    np.testing.assert_allclose(Fx, o["Fx"], atol=ATOL)
    np.testing.assert_allclose(Fy, o["Fy"], atol=ATOL)
    np.testing.assert_allclose(Fz, o["Fz"], atol=ATOL)
    np.testing.assert_allclose(Mr, o["M_real"], atol=ATOL)


def test_eigenvalues_golden():
    g = G["test_flux_eigenvalues_golden"]["eigenvalues"]
    o = g["output"]
    vmin, vmax, _, _ = O.eigenvalues6(g["input"], 1, 1, 0.5)
    assert abs(vmin - o["v6min_x"]) < 1e-7 and abs(vmax - o["v6max_x"]) < 1e-7
    vmin, vmax, _, _ = O.eigenvalues6(g["input"], 2, 1, 0.5)
    assert abs(vmin - o["v6min_y"]) < 1e-7 and abs(vmax - o["v6max_y"]) < 1e-7


@pytest.mark.parametrize("case", ["S2_case1", "S2_case2", "S2_case3"])
def test_realizability_S2(case):
    g = G["test_realizability_golden"][case]
    i, o = g["input"], g["output"]
    r = O.realizability_S2(i["S110"], i["S101"], i["S011"])
    np.testing.assert_allclose(r, [o["S110r"], o["S101r"], o["S011r"], o["S2r"]], atol=ATOL)


def test_check2d():
    g = G["test_diagnostics_golden"]["check2d"]
    names = ["S30", "S40", "S11", "S21", "S31", "S12", "S22", "S03", "S13", "S04"]
    out = O.check2D([g["input"][n] for n in names])
    np.testing.assert_allclose(out, [g["output"][n] for n in names], atol=ATOL)


def test_collision_basic():
    g = G["test_collision_golden"]["basic"]
    i = g["input"]
    out = O.collision35(i["M"], i["dt"], i["Kn"])
    np.testing.assert_allclose(out, g["output"], atol=ATOL)


def test_pas_hll_identity():
    g = G["test_numerical_schemes_golden"]["pas_hll"]
    i = g["input"]
    out = O.pas_hll(i["M"], i["F"], i["dt"], i["dx"], i["vpmin"], i["vpmax"])
    np.testing.assert_allclose(out, g["output"], atol=ATOL)
