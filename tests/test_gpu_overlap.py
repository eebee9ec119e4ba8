"""The fast modes overlap the pass-2 list kernel of step n with the pass 1 of step n+1 (a
second stream) inside hyq_run; hyq_step joins that stream before it returns, so a run of K
single steps takes the same kernels without any overlap.  Both must give the same states:
bitwise in mixed_newton (the cells of the pass-2 list take the same fast pass 1, only later),
and to the root-iteration tolerance in mixed_learned (those cells then start from the
analytic bounds instead of the network's seeds).  The crossing jets at 24x24x12, Ma=1.5,
Kn=0.01 stay finite for 9 steps in the oracle (the 10th meets a non-finite z Jacobian), so
the runs take 6."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _params(precision):
    import hyqmom_amd as H
    bg, jets = H.crossing_jets(1.5)
    return dict(Nx=24, Ny=24, Nz=12, tmax=1e9, Kn=0.01, Ma=1.5, flag2D=0, CFL=0.7, Nmom=35,
                nnmax=10 ** 9, dtmax=1e-2, rhol=1.0, rhor=0.1, T=1.0, r110=0.0, r101=0.0, r011=0.0,
                symmetry_check_interval=0, homogeneous_z=False, debug_output=False,
                use_custom_ic=True, ic_background=bg, ic_jets=jets, precision=precision)


def _two_ways(precision, steps, model=None):
    import hyqmom_amd as H
    p = _params(precision)
    if model is not None:
        p["warmstart_model"] = model
    M0 = H.initial_state(p)
    a = H.Solver(p)
    a.set_state(M0)
    a.run(1e9, steps)
    ra = (a.get_state(), a.t, a.nn, a.deferred_cells())
    a.close()
    b = H.Solver(p)
    b.set_state(M0)
    for _ in range(steps):
        b.step()
    rb = (b.get_state(), b.t, b.nn)
    b.close()
    return ra, rb


def test_newton_run_equals_single_steps_bitwise():
    (Ma, ta, na, dfr), (Mb, tb, nb) = _two_ways("mixed_newton", 6)
    print("deferred cells in the last step (pass 1, pass 2):", dfr)
    assert na == nb == 6 and ta == tb
    assert np.array_equal(Ma.view(np.int64), Mb.view(np.int64))


def test_learned_run_matches_single_steps():
    import mlp_ref as R
    from hyqmom_amd.warmstart import WarmStartMLP
    (Ma, ta, na, _), (Mb, tb, nb) = _two_ways("mixed_learned", 6, WarmStartMLP(R.random_blob(3)))
    assert na == nb == 6
    assert abs(ta - tb) <= 1e-12 * tb
    scale = np.abs(Mb).max(axis=(0, 1, 2))
    assert (np.abs(Ma - Mb) / scale).max() < 1e-9
