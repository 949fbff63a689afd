"""Pins of the Kalman filter oracle (oracle/kalman.py) on the CPU.

* the update (kalman_filter.py:196-233, explicit inverse) against the same
  posterior written through the information matrix (Woodbury identity), for
  the filter of type 0 (measurement noise 1/eta after its growth step) and of
  type 1 (noise lambda, fading memory 1/lambda), and against the one-weight
  closed form;
* the observation model (scaled descriptors + centre-only descriptor
  Jacobians from the C oracle, numpy MLP in forward mode) against the C
  oracle's energy and self-term forces on the reference's h2o fixture;
* complex-step Jacobians against central differences of the same model;
* a planted mistake (a dropped q I term) is caught by the information form.
"""

import random

import numpy as np
import pytest

from conftest import h2o_fixture_inputs
from oracle import kalman as okf


def _spd(rng, n):
    B = rng.normal(size=(n, n)) * 0.3
    return B @ B.T + np.eye(n) * 2.0


@pytest.mark.parametrize("kalman_type,m", [(0, 1), (0, 5), (1, 1), (1, 7)])
def test_update_equals_information_form(kalman_type, m):
    rng = np.random.default_rng(10 + m + 7 * kalman_type)
    n_w = 12
    st = okf.FilterState(rng.normal(size=n_w), kalman_type=kalman_type)
    st.P = _spd(rng, n_w)
    P0, W0 = st.P.copy(), st.W[:, 0].copy()
    H = rng.normal(size=(n_w, m))
    xi = rng.normal(size=m)
    eta0, lam0, q0 = st.eta, st.lambda0, st.q
    okf.ekf_update(st, xi, H)
    if kalman_type == 0:
        eta = eta0 * np.exp(0.01) if eta0 < 1.0 else eta0
        r, c = 1.0 / eta, 1.0
        assert st.eta == pytest.approx(eta, rel=1e-15)
    else:
        r, c = lam0, 1.0 / lam0
        assert st.lambda0 == pytest.approx(0.9995 * lam0 + 1 - 0.9995, rel=1e-15)
    P_want, W_want = okf.information_form(P0, W0, H, xi, r, c, q0)
    np.testing.assert_allclose(st.P, P_want, rtol=0, atol=1e-11 * np.abs(P_want).max())
    np.testing.assert_allclose(st.W[:, 0], W_want, rtol=0, atol=1e-11 * np.abs(W_want - W0).max())
    np.testing.assert_allclose(st.P, st.P.T, rtol=0, atol=1e-12 * np.abs(P_want).max())
    assert st.q == pytest.approx(q0 * np.exp(-2.302), rel=1e-15)


def test_planted_missing_process_noise_is_caught():
    rng = np.random.default_rng(3)
    n_w, m = 8, 3
    st = okf.FilterState(rng.normal(size=n_w))
    st.P = _spd(rng, n_w)
    P0, W0 = st.P.copy(), st.W[:, 0].copy()
    H, xi = rng.normal(size=(n_w, m)), rng.normal(size=m)
    q0 = st.q
    okf.ekf_update(st, xi, H)
    st.P -= q0 * np.eye(n_w)   # the mistake
    P_want, _ = okf.information_form(P0, W0, H, xi, 1.0 / (0.01 * np.exp(0.01)), 1.0, q0)
    assert np.abs(st.P - P_want).max() > 1e-4


def test_one_weight_closed_form():
    st = okf.FilterState(np.array([0.3]))
    st.P = np.array([[2.5]])
    okf.ekf_update(st, np.array([0.7]), np.array([[1.9]]))
    r = 1.0 / (0.01 * np.exp(0.01))
    p, w = okf.ekf_update_scalar(2.5, 0.3, 1.9, 0.7, r, 0.01)
    assert st.P[0, 0] == pytest.approx(p, rel=1e-14)
    assert st.W[0, 0] == pytest.approx(w, rel=1e-14)


def test_reference_random_generators():
    """The fit loop draws its order and energy/force choices as kalman_filter.py:75-76, 126-148 do:
    global random/np.random seeded once, one shuffle per epoch, one rand() per structure."""
    seed = 1234567
    random.seed(seed)
    np.random.seed(seed)
    idx = list(range(9))
    want = []
    for _ in range(3):
        random.shuffle(idx)
        want.append((list(idx), [np.random.rand() < 0.3 for _ in idx]))
    a, b = random.Random(seed), np.random.RandomState(seed)
    idx = list(range(9))
    for order, choices in want:
        a.shuffle(idx)
        assert idx == order
        assert [b.rand() < 0.3 for _ in idx] == choices


def test_observation_model_matches_c_oracle():
    pos, types, L, nnp = h2o_fixture_inputs(np.float64)
    obs = okf.Observations(nnp, pos, types, L)
    E, F = obs.energy_forces(nnp.layers)
    E_c, _, f_self, _ = nnp.energy_forces(pos, types, L)
    assert E == pytest.approx(E_c, rel=1e-13)
    np.testing.assert_allclose(F, f_self, rtol=0, atol=1e-12 * np.abs(f_self).max())


def test_complex_step_jacobians_match_central_differences():
    pos, types, L, nnp = h2o_fixture_inputs(np.float64)
    obs = okf.Observations(nnp, pos, types, L)
    els = nnp.elements
    w, _ = okf.flatten_layers(nnp.layers, els)
    E_ref, F_ref = -10.0, np.zeros((len(types), 3))
    xi_E, H_E, xi_F, H_F = okf.errors_and_jacobians(obs, w, nnp.layers, els, E_ref, F_ref)
    n = len(types)
    rng = np.random.default_rng(0)
    h = 1e-6
    for p in rng.choice(w.size, size=12, replace=False):
        vals = []
        for s in (1.0, -1.0):
            wp = w.copy()
            wp[p] += s * h
            E, F = obs.energy_forces(okf.unflatten_layers(wp, nnp.layers, els))
            vals.append(((E_ref - E) / n, (F_ref - F).ravel()))
        dE = -(vals[0][0] - vals[1][0]) / (2 * h)
        dF = -(vals[0][1] - vals[1][1]) / (2 * h)
        assert H_E[p] == pytest.approx(dE, rel=1e-6, abs=1e-9)
        np.testing.assert_allclose(H_F[p], dF, rtol=0, atol=1e-7 * max(np.abs(dF).max(), 1e-3))
    assert np.abs(H_F).max() > 0 and np.abs(H_E).max() > 0
