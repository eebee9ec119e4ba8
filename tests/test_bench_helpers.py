"""bench.py's host-side helpers: the seeded one-ulp perturbations of the initial state that
replace a trajectory that breaks down, and the algorithmic bytes of the update."""
import numpy as np

import bench


def test_perturbation_is_seeded_and_one_ulp():
    M0 = np.asfortranarray(np.random.default_rng(0).random((12, 10, 6, 35)) + 0.5)
    assert bench.perturbed_state(M0, 0) is M0
    a, b = bench.perturbed_state(M0, 3), bench.perturbed_state(M0, 3)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, bench.perturbed_state(M0, 4))
    d = a != M0
    assert not d[..., 1:].any()  # only densities move
    n = int(d[..., 0].sum())
    assert 0 < n <= 100
    assert np.array_equal(a[d], np.nextafter(M0[d], np.inf))


def test_update_bytes():
    # read Mr (35 FP64), three flux blocks (3 x 35) and 6 speeds, write the state (35 FP64)
    assert bench.update_bytes("fp64") == 280 + 840 + 48 + 280
    assert bench.update_bytes("mixed_learned") == 280 + 420 + 24 + 280


def test_measured_traffic_lookup(tmp_path, monkeypatch):
    f = tmp_path / "t.json"
    f.write_text('{"runs": {"mixed_learned 120x120x60 20 steps": {"pass1": 2.5e9}}}')
    monkeypatch.setattr(bench, "TRAFFIC_FILE", str(f))
    assert bench.measured_traffic("mixed_learned", (120, 120, 60), 20, "pass1") == 2.5e9
    assert bench.measured_traffic("mixed_learned", (120, 120, 60), 200, "pass1") is None
    assert bench.measured_traffic("fp64", (120, 120, 60), 20, "pass1") is None
    assert bench.measured_traffic("mixed_learned", (120, 120, 60), 20, "update") is None


def test_cpu_baseline_record():
    import hyqmom_amd as H
    M = np.asfortranarray(H.initial_state(bench.jets_params((40, 40, 40), "fp64")))
    r = bench.cpu_baseline(M, 0, seconds=0.2)
    assert r["value"] > 0 and r["one_thread"]["value"] > 0
    assert r["mpi_ranks"] == 1 and r["kind"] == "port" and r["cores"] >= 1
    assert "16^3" in r["sample"] and "8^3" in r["one_thread"]["sample"]
