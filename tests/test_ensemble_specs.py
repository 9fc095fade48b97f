"""Host-side checks of the batching workload's generator (no GPU)."""
import numpy as np

from flow_amd import problems


def test_specs_are_seeded_and_in_range():
    a = problems.ensemble_specs(count=64)
    b = problems.ensemble_specs(count=64)
    assert a == b
    for n, kind, split, ratio in a:
        assert 64 <= n <= 1024
        assert 0.3 <= split <= 0.7
        assert 10.0 <= ratio <= 1000.0
        assert kind in ("sod_x", "sod_y", "blast")


def test_instance_states_are_interior_only_and_positive():
    for spec in problems.ensemble_specs(count=8, n_lo=8, n_hi=40):
        rho, e = problems.ensemble_instance(*spec)
        n = spec[0]
        L = n + 4
        R = rho[: L * L].reshape(L, L)
        E = e[: L * L].reshape(L, L)
        assert np.all(R[2:-2, 2:-2] > 0) and np.all(E[2:-2, 2:-2] > 0)
        assert np.all(R[:2] == 0) and np.all(R[:, :2] == 0)
        assert np.all(rho[L * L:] == 0)
