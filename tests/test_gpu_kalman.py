"""Kalman filter trainer on the GPU against the oracle (oracle/kalman.py).

* pt_mlp_jacobian (force rows) and pt_mlp_train_grad (energy row) against
  complex-step Jacobians of the oracle's observation model, on the
  reference's h2o fixture (potential h2o.json + weights, structures of
  h2o.data): FP64, 1e-10 of the largest entry;
* pt_kalman_update (rocBLAS + rocSOLVER Cholesky) against the oracle's
  explicit-inverse update, filter types 0 and 1, m = 1 to 1,800: m <= n_w
  takes the m x m factorisation, m > n_w the n_w x n_w one;
* KalmanFilter.fit over h2o.data (force fraction 0.5, three epochs, types 0
  and 1) against the oracle's fit loop with the reference's random
  generators: weights and losses;
* one force update at 1,536 atoms (m = 4,608 observations, the student
  network) against the information form of the same posterior, computed in
  FP64 with torch.linalg as the checker.
"""

import ctypes
import time

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import kalman as okf
from oracle.bridge import oracle_nnp

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _fixture(kalman_type=0, force_fraction=0.5, epochs=3):
    from pantea_amd.io import Dataset
    from pantea_amd.potential import NeuralNetworkPotential

    pot = NeuralNetworkPotential.from_runner(GOLDEN / "h2o.json")
    pot.load()
    st = pot.settings
    st.kalman_type = kalman_type
    st.short_force_fraction = force_fraction
    st.epochs = epochs
    st.updater_type = "kalman_filter"
    ds = Dataset.from_runner(GOLDEN / "h2o.data")
    return pot, ds


def _oracle_obs(pot, s):
    orc = oracle_nnp(pot, s)
    pos = s.positions.cpu().numpy()
    obs = okf.Observations(orc, pos, s.atom_types_host, np.asarray(s.box_lengths))
    return orc, obs


def test_jacobians_match_complex_step():
    from pantea_amd.kalman import KalmanFilter

    pot, ds = _fixture()
    kf = KalmanFilter.from_settings(pot, pot.settings)
    kf.dev = torch.device("cuda")
    kf._bind()
    items = kf._prepare(ds)
    kf._buffers(items)
    for k, it in enumerate(items):
        s = ds[k]
        orc, obs = _oracle_obs(pot, s)
        w, _ = okf.flatten_layers(orc.layers, orc.elements)
        np.testing.assert_array_equal(kf.W.cpu().numpy(), w)
        F_ref = s.forces.cpu().numpy()
        xi_E, H_E, xi_F, H_F = okf.errors_and_jacobians(obs, w, orc.layers, orc.elements, float(s.total_energy),
                                                        F_ref)
        xe, He = kf.calculate_loss_energy(it)
        xe, He = float(xe.item()), He.cpu().numpy()[0]
        print(f"structure {k}: xi_E {xe:.12e} vs {xi_E:.12e}, |H_E| {np.abs(H_E).max():.3e}", flush=True)
        assert xe == pytest.approx(xi_E, rel=1e-11, abs=1e-14)
        np.testing.assert_allclose(He, H_E, rtol=0, atol=1e-10 * np.abs(H_E).max())
        xf, Hf = kf.calculate_loss_force(it)
        xf, Hf = xf.cpu().numpy(), Hf.cpu().numpy()   # rows 3 i + a, internal order
        perm = it.nl.perm.cpu().numpy()
        cols = (3 * perm[:, None] + np.arange(3)[None]).ravel()
        np.testing.assert_allclose(xf, xi_F[cols], rtol=0, atol=1e-11 * np.abs(xi_F).max())
        np.testing.assert_allclose(Hf, H_F[:, cols].T, rtol=0, atol=1e-10 * np.abs(H_F).max())


@pytest.mark.parametrize("kalman_type", [0, 1])
@pytest.mark.parametrize("n_w,m", [(40, 1), (40, 36), (40, 40), (40, 90), (600, 1800)])
def test_update_matches_oracle(kalman_type, n_w, m):
    from pantea_amd import _lib

    lib = _lib.load()
    rng = np.random.default_rng(n_w + m + kalman_type)
    st = okf.FilterState(rng.normal(size=n_w), kalman_type=kalman_type)
    B = rng.normal(size=(n_w, n_w)) * 0.05
    st.P = np.eye(n_w) * 50.0 + B @ B.T
    H = rng.normal(size=(n_w, m)) * 0.02
    xi = rng.normal(size=m)
    P_d = torch.as_tensor(st.P.T.copy(), device="cuda")
    W_d = torch.as_tensor(st.W[:, 0].copy(), device="cuda")
    H_d = torch.as_tensor(H.T.copy(), device="cuda")
    x_d = torch.as_tensor(xi, device="cuda")
    if kalman_type == 0:
        eta = st.eta * np.exp(st.eta_tau)
        r, c = 1.0 / eta, 1.0
    else:
        r, c = st.lambda0, 1.0 / st.lambda0
    q = st.q
    okf.ekf_update(st, xi, H)
    ws = torch.empty(lib.pt_kalman_workspace_bytes(n_w, m), dtype=torch.uint8, device="cuda")
    info = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    _lib.check(lib.pt_kalman_update(n_w, m, P_d.data_ptr(), H_d.data_ptr(), n_w, x_d.data_ptr(), W_d.data_ptr(),
                                    r, c, q, ws.data_ptr(), ws.numel(), info.data_ptr(), _lib.stream_ptr()),
               "pt_kalman_update")
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    P_g = P_d.cpu().numpy().T
    W_g = W_d.cpu().numpy()
    np.testing.assert_allclose(P_g, st.P, rtol=0, atol=1e-10 * np.abs(st.P).max())
    dW = st.W[:, 0] - W_g
    assert np.abs(dW).max() <= 1e-10 * max(np.abs(st.W[:, 0]).max(), 1.0)


@pytest.mark.parametrize("n_w,m,col", [(40, 90, 1), (130, 200, 70), (40, 36, 1)])
def test_update_reports_a_failed_factorisation(n_w, m, col):
    """An indefinite matrix to factorise sets info to the 1-based column of the first
    pivot that is not positive, as dpotrf reports it (P itself in the n_w x n_w form,
    A = H^T P H + r I in the m x m form)."""
    from pantea_amd import _lib

    lib = _lib.load()
    rng = np.random.default_rng(7)
    P = np.eye(n_w) * 50.0
    if m > n_w:
        P[col - 1, col - 1] = -1.0          # P is indefinite from this column on
    else:
        P = -np.eye(n_w)                    # H^T P H + r I is negative definite
    H = rng.normal(size=(n_w, m)) * (0.02 if m > n_w else 1.0)
    P_d = torch.as_tensor(P.T.copy(), device="cuda")
    W_d = torch.zeros(n_w, dtype=torch.float64, device="cuda")
    H_d = torch.as_tensor(H.T.copy(), device="cuda")
    x_d = torch.as_tensor(rng.normal(size=m), device="cuda")
    ws = torch.empty(lib.pt_kalman_workspace_bytes(n_w, m), dtype=torch.uint8, device="cuda")
    info = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    _lib.check(lib.pt_kalman_update(n_w, m, P_d.data_ptr(), H_d.data_ptr(), n_w, x_d.data_ptr(), W_d.data_ptr(),
                                    0.01, 1.0, 0.0, ws.data_ptr(), ws.numel(), info.data_ptr(), _lib.stream_ptr()),
               "pt_kalman_update")
    torch.cuda.synchronize()
    assert int(info.item()) == col


@pytest.mark.parametrize("kalman_type", [0, 1])
def test_fit_matches_oracle_fit(kalman_type):
    from pantea_amd.kalman import NeuralNetworkPotentialTrainer

    pot, ds = _fixture(kalman_type=kalman_type, force_fraction=0.5, epochs=3)
    structs = []
    orc0 = None
    for k in range(len(ds)):
        s = ds[k]
        orc, obs = _oracle_obs(pot, s)
        orc0 = orc0 or orc
        structs.append((obs, float(s.total_energy), s.forces.cpu().numpy()))
    w0, _ = okf.flatten_layers(orc0.layers, orc0.elements)
    st = okf.FilterState(w0, kalman_type=kalman_type)
    hist_o = okf.ekf_fit(st, structs, orc0.layers, orc0.elements, force_fraction=0.5, epochs=3,
                         seed=pot.settings.random_seed)
    tr = NeuralNetworkPotentialTrainer.from_settings(pot, pot.settings)
    hist = tr.fit_model(ds, log=print)
    W = tr.updater.W.cpu().numpy()
    dw = np.abs(W - st.W[:, 0]).max()
    moved = np.abs(st.W[:, 0] - w0).max()
    print(f"type {kalman_type}: weights moved {moved:.3e}, |W - W_oracle| {dw:.3e}", flush=True)
    assert moved > 1e-6
    assert dw <= 1e-9 * max(np.abs(w0).max(), 1.0)
    np.testing.assert_allclose(hist["loss_energy"], hist_o["loss_energy"], rtol=1e-8, atol=1e-14)
    np.testing.assert_allclose(hist["loss_force"], hist_o["loss_force"], rtol=1e-8, atol=1e-14)
    P = tr.updater.P.cpu().numpy()
    np.testing.assert_allclose(P, st.P, rtol=0, atol=1e-9 * np.abs(st.P).max())
    # the potential carries the trained weights
    flat, _ = okf.flatten_layers(oracle_nnp(pot, ds[0]).layers, pot.elements)
    np.testing.assert_array_equal(flat, W)


def test_force_update_at_1536_atoms_equals_information_form():
    from pantea_amd import bench_models, water
    from pantea_amd.atoms import Structure
    from pantea_amd.kalman import KalmanFilter

    from pantea_amd.potential import NNPSettings

    # 512 molecules at the density of the 32,768-molecule box (187.7 Bohr)
    s = Structure.from_dict(water.water_box(512, 187.7 / 4, seed=5))
    student = bench_models.water_student(None, seed=12, precision="float64")
    student.fit_scaler([s])
    E, F = student.energy_and_forces(s, mode="reference", precision="float64")
    # labels: the student's own forces plus a smooth perturbation, energy shifted
    s.forces = F + 0.01 * torch.sin(s.positions)
    s.total_energy = E + 0.5
    settings = NNPSettings(elements=list(student.elements), global_nodes_short=[16, 16],
                           global_activation_short=["tanh", "tanh", "identity"], symfunction_short=[])
    kf = KalmanFilter.from_settings(student, settings)
    kf.dev = torch.device("cuda")
    kf._bind()
    items = kf._prepare([s])
    kf._buffers(items)
    P0 = kf.P.clone()
    W0 = kf.W.clone()
    xi, H = kf.calculate_loss_force(items[0])
    xi, H = xi.clone(), H.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kf.update_state_vector(xi, H)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    n_w, m = kf.num_states, xi.numel()
    print(f"force update: n_w {n_w}, m {m}: {dt * 1e3:.1f} ms", flush=True)
    assert m == 3 * s.natoms == 4608
    r = 1.0 / kf.eta
    Hm = H.T.contiguous()                      # n_w x m
    Minv = torch.linalg.inv(torch.linalg.inv(P0) + Hm @ Hm.T / r)
    P_want = Minv + 0.01 * torch.eye(n_w, dtype=torch.float64, device="cuda")
    W_want = W0 + Minv @ (Hm @ xi) / r
    assert int(kf.info_sum.item()) == 0
    err_P = float((kf.P - P_want).abs().max() / P_want.abs().max())
    err_W = float((kf.W - W_want).abs().max() / (W_want - W0).abs().max())
    print(f"|dP| {err_P:.2e} |dW| {err_W:.2e}", flush=True)
    assert err_P <= 1e-9 and err_W <= 1e-8
