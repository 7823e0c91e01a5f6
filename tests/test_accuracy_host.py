"""Error measures and the round-off-sized perturbation of wamp/accuracy.py, on the CPU."""
import numpy as np

from wamp import accuracy, cases


def test_compare_of_a_state_with_itself_is_zero():
    cfg, U0 = cases.triple_point(64, 32)
    err = accuracy.compare(cfg, U0, U0)
    assert err["rho"] == err["p"] == err["alpha1"] == err["centroid_shift_cells"] == 0.0
    assert err["finite"]


def test_relative_l1_is_scale_free_and_known():
    a = np.array([1.0, 2.0, 3.0])
    assert abs(accuracy.rel_l1(a * (1 + 1e-3), a) - 1e-3) < 1e-15
    assert accuracy.rel_l1(2 * a, a) == 1.0


def test_perturbation_is_seeded_interior_only_and_eps_sized():
    cfg, U0 = cases.triple_point(64, 32)
    eps = 2.0 ** -24
    U1 = accuracy.perturbed(cfg, U0, eps, seed=3)
    U2 = accuracy.perturbed(cfg, U0, eps, seed=3)
    assert np.array_equal(U1, U2)
    assert not np.array_equal(U1, accuracy.perturbed(cfg, U0, eps, seed=4))
    si, sj = cfg.grid.interior()
    inner = np.zeros(U0.shape[1:], dtype=bool)
    inner[si, sj] = True
    for m in range(6):
        if m in (0, 1, 4):
            r = (U1[m] - U0[m])[inner] / U0[m][inner]
            assert np.allclose(np.abs(r), eps, rtol=1e-6)  # every interior cell moved by +-eps
            assert 0.4 < (r > 0).mean() < 0.6
        else:
            assert np.array_equal(U1[m], U0[m])  # momenta and alpha1 untouched
        assert np.array_equal(U1[m][~inner], U0[m][~inner])  # ghosts untouched
    err = accuracy.compare(cfg, U1, U0)
    assert 0.5 * eps < err["rho"] < 1.5 * eps and err["alpha1"] == 0.0
