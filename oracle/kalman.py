"""FP64 CPU oracle of pantea's global extended Kalman filter trainer --
TEST INFRASTRUCTURE ONLY (imported by tests/ as the checker; the product
package never imports it).

Restates potentials/nnp/kalman_filter.py:

  energy observation   xi = (E* - E) / N, H = -dxi/dW          :235-249, :276-290
  force observation    xi = F* - F_self (3N), H = -dxi/dW      :251-273, :293-335
  update               X = P H, A = H^T X + (1/eta | lambda) I,
                       K = X inv(A), P -= K X^T, P /= lambda (type 1),
                       P += q I, W += K xi, schedules           :196-233
  fit loop             random.shuffle + np.random.rand() < force_fraction
                       per structure, seeds from random_seed     :75-76, :112-194

The observation model takes its descriptors and centre-only descriptor
Jacobians from the C oracle (pyoracle.acsf / acsf_selfjac) and evaluates the
MLP and its input-gradient projection in numpy; the Jacobians dxi/dW come
from complex-step differentiation (Im f(W + i h e_p) / h, exact to rounding
for analytic activations), not from a restated derivative.
"""

from __future__ import annotations

import random
from typing import Dict, List, Sequence, Tuple

import numpy as np

from . import pyoracle

# ---------------------------------------------------------------------------
# flat state vector: elements in potential order, per layer kernel (in x out,
# row-major) then bias
# ---------------------------------------------------------------------------


def flatten_layers(layers: Dict[str, list], elements: Sequence[str]) -> Tuple[np.ndarray, Dict[str, int]]:
    parts, offs, o = [], {}, 0
    for e in elements:
        offs[e] = o
        for W, b, _ in layers[e]:
            parts.append(np.asarray(W, np.float64).ravel())
            parts.append(np.asarray(b, np.float64).ravel())
            o += W.size + b.size
    return np.concatenate(parts), offs


def unflatten_layers(w: np.ndarray, template: Dict[str, list], elements: Sequence[str]) -> Dict[str, list]:
    out, o = {}, 0
    for e in elements:
        lay = []
        for W, b, a in template[e]:
            Wn = w[o:o + W.size].reshape(W.shape)
            o += W.size
            bn = w[o:o + b.size]
            o += b.size
            lay.append((Wn, bn, a))
        out[e] = lay
    return out


def _act_c(name: str, z):
    """Activation and derivative for real or complex z (analytic forms)."""
    if name == "identity":
        return z, np.ones_like(z)
    if name == "tanh":
        t = np.tanh(z)
        return t, 1.0 - t * t
    if name == "logistic":
        s = 1.0 / (1.0 + np.exp(-z))
        return s, s * (1.0 - s)
    if name == "gaussian":
        g = np.exp(-0.5 * z * z)
        return g, -z * g
    if name == "cos":
        return np.cos(z), -np.sin(z)
    if name == "exp":
        e = np.exp(-z)
        return e, -e
    if name == "harmonic":
        return z * z, 2.0 * z
    if name == "softplus":
        return np.log(1.0 + np.exp(z)), 1.0 / (1.0 + np.exp(-z))
    raise ValueError(f"activation {name} is not analytic")


def mlp_value_and_projections(layers, x, t):
    """y = MLP(x) (n,) and ydot[d] = grad_x y . t[d] (n_dir, n) by forward mode
    (model.py:24-58); x: (n, k), t: (n_dir, n, k); real or complex."""
    h = x
    hd = t
    for W, b, a in layers:
        z = h @ W + b
        zd = hd @ W
        h, d = _act_c(a, z)
        hd = zd * d[None]
    return h[:, 0], hd[:, :, 0]


class Observations:
    """Energy and self-term forces of one structure as functions of W."""

    def __init__(self, orc: pyoracle.OracleNNP, pos, types, box):
        self.orc = orc
        self.n = len(types)
        self.types = np.asarray(types)
        G = orc.descriptors(pos, types, box)
        self.x, self.t, self.idx = {}, {}, {}
        for e in orc.elements:
            idx = np.nonzero(self.types == orc.type_of[e])[0]
            self.idx[e] = idx
            if idx.size == 0:
                continue
            J = pyoracle.acsf_selfjac(pos, types, box, orc.tables[e], center_pos=np.asarray(pos)[idx],
                                      self_idx=idx.astype(np.int32))  # (n_e, n_sf, 3)
            slope = pyoracle.scaler_slope(orc.scaler_kind, orc.scaler_params[e], orc.smin, orc.smax)
            self.x[e] = pyoracle.scaler_transform(orc.scaler_kind, orc.scaler_params[e], G[e], orc.smin, orc.smax)
            self.t[e] = np.transpose(J, (2, 0, 1)) * slope  # (3, n_e, n_sf), scaled

    def energy_forces(self, layers) -> Tuple[complex, np.ndarray]:
        """E (total, energy.py:45-63) and F_self (n, 3, force.py:28-43) for the given layers."""
        E = 0.0
        F = None
        for e in self.orc.elements:
            idx = self.idx[e]
            if idx.size == 0:
                continue
            y, yd = mlp_value_and_projections(layers[e], self.x[e], self.t[e])
            E = E + y.sum()
            if F is None:
                F = np.zeros((self.n, 3), dtype=np.result_type(y, np.float64))
            F[idx] = -yd.T
        return E, F


def errors_and_jacobians(obs: Observations, w: np.ndarray, template, elements, E_ref: float, F_ref: np.ndarray,
                         h: float = 1e-30):
    """xi_E, H_E (n_w,), xi_F (3N,), H_F (n_w, 3N): the errors of kalman_filter.py:276-303 and
    H = -d xi / dW by complex-step differentiation over every weight."""
    n = obs.n
    E, F = obs.energy_forces(unflatten_layers(w, template, elements))
    xi_E = (E_ref - E) / n
    xi_F = (np.asarray(F_ref) - F).ravel()
    n_w = w.size
    H_E = np.zeros(n_w)
    H_F = np.zeros((n_w, 3 * n))
    for p in range(n_w):
        wc = w.astype(np.complex128)
        wc[p] += 1j * h
        Ec, Fc = obs.energy_forces(unflatten_layers(wc, template, elements))
        H_E[p] = np.imag(Ec) / h / n          # -d((E* - E)/N)/dW
        H_F[p] = np.imag(Fc).ravel() / h      # -d(F* - F)/dW
    return float(xi_E), H_E, xi_F, H_F


# ---------------------------------------------------------------------------
# the filter (kalman_filter.py:77-110, 196-233)
# ---------------------------------------------------------------------------


class FilterState:
    def __init__(self, w: np.ndarray, kalman_type=0, epsilon=0.01, q0=0.01, qtau=2.302, qmin=1e-6, eta=0.01,
                 etamax=1.0, lambda0=0.96, neu=0.9995):
        self.W = np.asarray(w, np.float64).reshape(-1, 1).copy()
        self.P = np.eye(self.W.shape[0]) / epsilon
        self.kalman_type = kalman_type
        self.q, self.q_tau, self.q_min = q0, qtau, qmin
        # eta_tau starts from kalman_eta (kalman_filter.py:98)
        self.eta, self.eta_tau, self.eta_max = eta, eta, etamax
        self.lambda0, self.neu, self.gamma = lambda0, neu, 1.0


def ekf_update(st: FilterState, Xi: np.ndarray, H: np.ndarray) -> None:
    Xi = np.asarray(Xi, np.float64).reshape(-1, 1)
    H = np.asarray(H, np.float64).reshape(st.W.shape[0], -1)
    X = st.P @ H
    A = H.T @ X
    if st.kalman_type == 0 and st.eta < st.eta_max:
        st.eta *= np.exp(st.eta_tau)
    m = Xi.shape[0]
    if st.kalman_type == 0:
        A = A + (1.0 / st.eta) * np.eye(m)
    elif st.kalman_type == 1:
        A = A + st.lambda0 * np.eye(m)
    K = X @ np.linalg.inv(A)
    st.P = st.P - K @ X.T
    if st.kalman_type == 1:
        st.P = st.P * (1.0 / st.lambda0)
    st.P = st.P + st.q * np.eye(st.P.shape[0])
    st.W = st.W + K @ Xi
    if st.q > st.q_min:
        st.q *= np.exp(-st.q_tau)
    if st.kalman_type == 1:
        st.lambda0 = st.neu * st.lambda0 + 1.0 - st.neu
        st.gamma = 1.0 / (1.0 + st.lambda0 / st.gamma)


def ekf_fit(st: FilterState, structures: List[Tuple[Observations, float, np.ndarray]], template, elements,
            force_fraction: float, epochs: int, seed: int):
    """The loop of kalman_filter.py:112-194 with the reference's random generators."""
    shuffle_rng = random.Random(seed)
    choice_rng = np.random.RandomState(seed)
    indices = list(range(len(structures)))
    history = {"loss_energy": [], "loss_force": []}
    for _ in range(epochs):
        shuffle_rng.shuffle(indices)
        le = lf = 0.0
        ne = nf = 0
        for i in indices:
            obs, E_ref, F_ref = structures[i]
            xi_E, H_E, xi_F, H_F = errors_and_jacobians(obs, st.W[:, 0], template, elements, E_ref, F_ref)
            if choice_rng.rand() < force_fraction:
                lf += float(xi_F @ xi_F)
                nf += 1
                ekf_update(st, xi_F, H_F)
            else:
                le += xi_E * xi_E
                ne += 1
                ekf_update(st, np.array([xi_E]), H_E.reshape(-1, 1))
        history["loss_energy"].append(le / (ne if ne else 1))
        history["loss_force"].append(lf / (nf if nf else 1))
    return history


def information_form(P0: np.ndarray, W0: np.ndarray, H: np.ndarray, xi: np.ndarray, r: float, c: float,
                     q: float):
    """The same update written through the information matrix:
    P+ = c (P0^-1 + H H^T / r)^-1 + q I,  W+ = W0 + (P0^-1 + H H^T / r)^-1 H xi / r
    (Woodbury identity applied to K = P H (H^T P H + r I)^-1)."""
    Minv = np.linalg.inv(np.linalg.inv(P0) + H @ H.T / r)
    return c * Minv + q * np.eye(P0.shape[0]), W0 + Minv @ H @ xi / r


def ekf_update_scalar(p: float, w: float, h: float, xi: float, r: float, q: float) -> Tuple[float, float]:
    """One weight, one observation: K = p h / (h^2 p + r)."""
    k = p * h / (h * h * p + r)
    return p - k * h * p + q, w + k * xi


__all__ = ["flatten_layers", "unflatten_layers", "mlp_value_and_projections", "Observations",
           "errors_and_jacobians", "FilterState", "ekf_update", "ekf_fit", "information_form",
           "ekf_update_scalar"]
