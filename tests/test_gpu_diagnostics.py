"""The runner's diagnostics on the device (hyq_symmetry_maxdiff, hyq_vmax_breakdown) against
the CPU reference: the symmetry check of simulation_runner.jl:565-628 on the crossing jets and
on a perturbed state, and the first-step vmax breakdown (:404-461) on the speeds of a step in
each precision mode."""
import numpy as np
import pytest

import diagnostics_ref as D

pytestmark = pytest.mark.gpu


def _params(precision, N=(24, 24, 12), Ma=1.5, Kn=0.01):
    import hyqmom_amd as H
    bg, jets = H.crossing_jets(Ma)
    return dict(Nx=N[0], Ny=N[1], Nz=N[2], tmax=1e9, Kn=Kn, Ma=Ma, flag2D=0, CFL=0.7, Nmom=35,
                nnmax=10 ** 9, dtmax=1e-2, rhol=1.0, rhor=0.1, T=1.0, r110=0.0, r101=0.0, r011=0.0,
                symmetry_check_interval=2, homogeneous_z=False, debug_output=False,
                use_custom_ic=True, ic_background=bg, ic_jets=jets, precision=precision)


@pytest.mark.parametrize("N", [(24, 24, 12), (17, 23, 5)])
def test_symmetry_maxdiff_matches_reference(N):
    import hyqmom_amd as H
    p = _params("fp64", N)
    M0 = H.initial_state(p)
    s = H.Solver(p)
    try:
        s.set_state(M0)
        s.run(1e9, 3)
        M = s.get_state()
        got, per = s.symmetry_maxdiff()
        want, wper = D.symmetry_maxdiff(M)
        assert abs(got - want) <= 1e-12 * max(1.0, abs(want)), (got, want)
        assert np.allclose(per, wper, rtol=1e-12, atol=1e-15)
        # a perturbed state: every diagonal pair differs
        rng = np.random.default_rng(3)
        Mp = np.asfortranarray(M * (1.0 + 1e-3 * rng.standard_normal(M.shape)))
        s.set_state(Mp)
        got, per = s.symmetry_maxdiff()
        want, wper = D.symmetry_maxdiff(Mp)
        assert want > 1e-6
        assert abs(got - want) <= 1e-12 * want, (got, want)
        assert np.allclose(per, wper, rtol=1e-12, atol=1e-15)
    finally:
        s.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed_newton", "mixed_learned"])
def test_vmax_breakdown_matches_reference(precision):
    import hyqmom_amd as H
    p = _params(precision)
    s = H.Solver(p)
    try:
        s.set_state(H.initial_state(p))
        s.step()
        vdir, worst = s.vmax_breakdown()
        vp = s.wave_speeds().reshape(6, -1)
        wdir, wworst = D.vmax_breakdown(vp, p["Nx"], p["Ny"], p["Nz"])
        assert vdir == wdir, (vdir, wdir)
        assert worst == wworst, (worst, wworst)
        assert max(vdir) > 1.0
    finally:
        s.close()


def test_runner_prints_symmetry_and_first_step(capsys):
    import hyqmom_amd as H
    p = _params("fp64", (20, 20, 10), Ma=1.0, Kn=1.0)  # stays finite (the smoke case)
    p["nnmax"] = 4
    M, t, nn, _ = H.simulation_runner(p, verbose=True)
    out = capsys.readouterr().out
    assert nn == 4
    assert out.count("MaxDiff") == 3  # steps 1, 2 and 4 with interval 2
    assert "vx_max" in out and "fastest cell" in out
