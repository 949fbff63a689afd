"""Pins the FP64 oracle (oracle/) before anything is compared against it.

Each check is chosen so that a plausible mistake (dropped term, wrong sign or
index, missing halving, wrong MIC) fails one of them:
  * the reference's golden vectors (test_acsf.py, test_nnp.py, test_structure.py),
  * closed forms (cutoff functions, scaler statistics),
  * brute force (neighbour sets from periodic images; G3 and G9 by explicit
    pair loops, G9 also by a three-atom closed form),
  * every MLP activation against its closed form and finite differences,
  * finite differences (self-Jacobian, self-term forces, -grad E),
  * invariants (sum of -grad E = 0, rotation invariance).
"""

import itertools
import json
import math

import numpy as np
import pytest

from conftest import GOLDEN, h2o_fixture_inputs
from oracle import pyoracle as O

TANHU, TANH, COS = O.CUTOFF_CODES["tanhu"], O.CUTOFF_CODES["tanh"], O.CUTOFF_CODES["cos"]


def _small_water(n_mol=8, L=14.0, seed=5):
    from pantea_amd import water

    d = water.water_box(n_mol, L, seed=seed, d_min=3.5)
    types = np.array([2 if e == "O" else 1 for e in d["elements"]], dtype=np.int32)
    return d["positions"], types, np.array([L, L, L])


def _teacher_rows(rc=6.0, code=TANH, angular=True):
    rows = []
    for ej in (1, 2):
        for eta in np.geomspace(0.001, 0.5, 3):
            rows.append([2, ej, 0, eta, 0.0, 0, 0, rc, code, 0])
    if angular:
        for ej, ek in ((1, 1), (1, 2), (2, 2)):
            for eta in (0.001, 0.05):
                for zeta in (1.0, 4.0):
                    for lam in (1.0, -1.0):
                        rows.append([3, ej, ek, eta, 0.0, zeta, lam, rc, code, 0])
        rows.append([9, 1, 2, 0.01, 0.0, 2.0, 1.0, rc, code, 0])
    return np.array(rows, dtype=np.float64)


# ---------------------------------------------------------------------------
# golden vectors of the reference's own tests
# ---------------------------------------------------------------------------

def test_golden_acsf_without_pbc(goldens):
    g = goldens["acsf_without_pbc"]
    pos = np.array(g["positions"])
    rows = [[2, 1, 0, eta, rs, 0, 0, g["cutoff"][1], TANHU, 0] for eta, rs in g["g2"]]
    G = O.acsf(pos, [1, 1], None, rows)
    assert G.shape == tuple(g["expected_shape"])
    exp = np.tile(np.array(g["expected_row"]), (2, 1))
    assert np.allclose(G, exp, rtol=goldens["allclose"]["rtol"], atol=goldens["allclose"]["atol"])
    assert np.max(np.abs(G - exp)) < 1e-10  # the golden has 10 decimals


def test_golden_acsf_with_pbc(goldens):
    g = goldens["acsf_with_pbc"]
    L = np.array(g["lattice"])
    types = np.array([2 if e == "O" else 1 for e in g["elements"]], dtype=np.int32)
    pos = O.wrap_positions(np.array(g["positions"]), L)
    rc = g["cutoff"][1]
    rows = [[2, 1, 0, g["g2_H"]["eta"], 0.0, 0, 0, rc, TANHU, 0],
            [3, 1, 1, g["g3_HH"]["eta"], 0.0, g["g3_HH"]["zeta"], g["g3_HH"]["lambda0"], rc, TANHU, 0]]
    idx = np.nonzero(types == 2)[0]
    G = O.acsf(pos, types, L, rows, center_pos=pos[idx], self_idx=idx.astype(np.int32))
    assert G.shape == tuple(g["expected_shape"])
    assert np.max(np.abs(G[g["atom_index"]] - np.array(g["expected_row"]))) < 1e-10


def test_golden_nnp_outputs(goldens):
    """test_nnp.py:217-254: energy and self-term forces of structure 0 of h2o.data."""
    g = goldens["nnp_outputs"]
    pos, types, L, nnp = h2o_fixture_inputs()
    E, e_atom, f_self, f_full = nnp.energy_forces(pos, types, L)
    F = np.array(g["forces"])
    # forces: the reference's own tolerance (jnp.allclose defaults)
    assert np.allclose(f_self, F, rtol=1e-5, atol=1e-8)
    # -grad E is a different quantity (the reference drops the cross terms)
    assert np.max(np.abs(f_full - F)) > 1e-2
    # energy: the golden was evaluated in float32 (test_nnp.py:17); 12 atomic
    # energies of |E_i| <= 0.25 from float32 descriptors of magnitude ~1 carry
    # a rounding error of a few 1e-7, so the bound is 2.5e-7 (3% of |E|; any
    # dropped atom or descriptor term moves E by > 1e-3).
    assert abs(E - g["energy"]) < 2.5e-7
    assert np.max(np.abs(e_atom)) > 1e-2


def test_golden_masses(goldens):
    m = goldens["masses"]
    for e in ("H", "O", "Ne"):
        assert O.mass_au(e) == pytest.approx(m[e], rel=1e-15)


# ---------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------

def test_tanh_prefactor_is_coth1_cubed():
    assert O.tanh_prefactor() == pytest.approx(1.0 / math.tanh(1.0) ** 3, rel=1e-15)


@pytest.mark.parametrize("code,name", [(0, "hard"), (1, "cos"), (2, "tanhu"), (3, "tanh"), (4, "exp"), (5, "poly1"), (6, "poly2")])
def test_cutoff_closed_forms(code, name):
    rc = 6.0 if code not in (5, 6) else 1.0
    rs = np.linspace(0.05, 0.999, 9) * rc
    for r in rs:
        x = r / rc
        expect = {
            "hard": 1.0,
            "cos": 0.5 * (math.cos(math.pi * x) + 1.0),
            "tanhu": math.tanh(1 - x) ** 3,
            "tanh": (math.tanh(1 - x) / math.tanh(1.0)) ** 3,
            "exp": math.exp(1.0 - 1.0 / (1.0 - x * x)),
            "poly1": (2 * r - 3) * r * r + 1,
            "poly2": ((15 - 6 * r) * r - 10) * r ** 3 + 1,
        }[name]
        assert O.cutoff(code, rc, r) == pytest.approx(expect, rel=1e-13, abs=1e-15)
        h = 1e-6 * rc
        fd = (O.cutoff(code, rc, r + h) - O.cutoff(code, rc, r - h)) / (2 * h)
        assert O.cutoff_der(code, rc, r) == pytest.approx(fd, rel=1e-6, abs=1e-8)
    assert O.cutoff(code, rc, rc) == 0.0 and O.cutoff(code, rc, 2 * rc) == 0.0


def test_mic_and_wrap():
    assert O.mic(6.0, 10.0) == -4.0 and O.mic(-6.0, 10.0) == 4.0 and O.mic(5.0, 10.0) == 5.0
    assert O.wrap(-0.5, 10.0) == 9.5 and O.wrap(10.5, 10.0) == 0.5 and O.wrap(3.0, 10.0) == 3.0
    # remainder semantics: tiny negative wraps to L (rounded), as jnp.remainder
    assert O.wrap(-1e-17, 10.0) == 10.0


def test_scaler_statistics_closed_form():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(100, 7))
    p = O.scaler_fit(x[:10])
    for a, b in ((10, 37), (37, 38), (38, 100)):
        p = O.scaler_partial_fit(p, x[a:b])
    assert p["nsamples"] == 100
    assert np.allclose(p["mean"], x.mean(0), atol=1e-13)
    assert np.allclose(p["sigma"], x.std(0), atol=1e-12)  # population std (ddof 0)
    assert np.array_equal(p["minval"], x.min(0)) and np.array_equal(p["maxval"], x.max(0))
    s = O.scaler_transform("scale", p, x, 0.0, 1.0)
    assert np.allclose(s.min(0), 0.0, atol=1e-14) and np.allclose(s.max(0), 1.0, atol=1e-14)
    c = O.scaler_transform("scale_center_sigma", p, x, 0.0, 1.0)
    assert np.allclose(c.mean(0), 0.0, atol=1e-12) and np.allclose(c.std(0), 1.0, atol=1e-12)


def test_mlp_gradient_fd():
    rng = np.random.default_rng(1)
    layers = [(rng.normal(size=(5, 7)), rng.normal(size=7), "tanh"),
              (rng.normal(size=(7, 4)), rng.normal(size=4), "softplus"),
              (rng.normal(size=(4, 1)), rng.normal(size=1), "identity")]
    x = rng.normal(size=(3, 5))
    y, g = O.mlp_forward_backward(layers, x)
    h = 1e-6
    for k in range(5):
        e = np.zeros(5); e[k] = h
        fd = (O.mlp_forward_backward(layers, x + e)[0] - O.mlp_forward_backward(layers, x - e)[0]) / (2 * h)
        assert np.allclose(g[:, k], fd, rtol=1e-7, atol=1e-9)
    # y = x @ W + b layout: a linear one-layer model is x @ W + b exactly
    W = rng.normal(size=(5, 1)); b = rng.normal(size=1)
    y1, g1 = O.mlp_forward_backward([(W, b, "identity")], x)
    assert np.allclose(y1, (x @ W + b)[:, 0], rtol=1e-15) and np.allclose(g1, np.tile(W.T, (3, 1)))


# closed forms of models/nn/activation.py:7-60 (n2p2 definitions): f(z)
_ACT_CLOSED = {
    "identity": lambda z: z,
    "tanh": math.tanh,
    "logistic": lambda z: 1.0 / (1.0 + math.exp(-z)),
    "softplus": lambda z: math.log1p(math.exp(z)),
    "relu": lambda z: max(z, 0.0),
    "gaussian": lambda z: math.exp(-0.5 * z * z),
    "cos": math.cos,
    "exp": lambda z: math.exp(-z),
    "harmonic": lambda z: z * z,
}


@pytest.mark.parametrize("act", sorted(_ACT_CLOSED))
def test_activation_values_and_derivatives(act):
    """Every activation: value against its closed form, derivative against a
    central finite difference of that closed form (relu away from 0), then the
    same activation inside a hidden layer through the MLP input gradient."""
    zs = np.array([-2.3, -0.7, -0.05, 0.11, 0.9, 1.7])
    for z in zs:
        y, d = O._act(act, np.array([z]))
        assert y[0] == pytest.approx(_ACT_CLOSED[act](z), rel=1e-14, abs=1e-15)
        h = 1e-6
        fd = (_ACT_CLOSED[act](z + h) - _ACT_CLOSED[act](z - h)) / (2 * h)
        assert d[0] == pytest.approx(fd, rel=1e-7, abs=1e-9)
    rng = np.random.default_rng(sum(map(ord, act)))
    layers = [(rng.normal(size=(4, 6)) * 0.5, rng.normal(size=6) * 0.1, act),
              (rng.normal(size=(6, 1)), rng.normal(size=1), "identity")]
    x = rng.normal(size=(5, 4))
    _, g = O.mlp_forward_backward(layers, x)
    h = 1e-6
    for k in range(4):
        e = np.zeros(4); e[k] = h
        fd = (O.mlp_forward_backward(layers, x + e)[0] - O.mlp_forward_backward(layers, x - e)[0]) / (2 * h)
        assert np.allclose(g[:, k], fd, rtol=1e-6, atol=1e-8)


# ---------------------------------------------------------------------------
# brute force
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("L,rc,n", [(10.0, 3.0, 60), (10.0, 5.0, 40), (7.0, 6.0, 30), (20.0, 4.5, 200)])
def test_neighbors_equal_nearest_image_sets(L, rc, n):
    rng = np.random.default_rng(int(L * 10 + rc))
    pos = rng.random((n, 3)) * L
    counts, idx = O.neighbors(pos, [L, L, L], rc)
    rows = np.split(idx, np.cumsum(counts)[:-1])
    for i in range(n):
        d = pos[i] - pos
        best = np.full(n, np.inf)
        for s in itertools.product((-1, 0, 1), repeat=3):
            best = np.minimum(best, np.linalg.norm(d + np.array(s) * L, axis=1))
        want = set(np.nonzero((best <= rc) & (best > 0))[0])
        got = set(rows[i].tolist())
        near = set(np.nonzero(np.abs(best - rc) < 1e-12)[0])
        assert (got ^ want) <= near
        assert list(rows[i]) == sorted(rows[i])


def test_g3_halving_matches_unordered_pairs():
    """G3 with equal neighbour elements = sum over unordered pairs (acsf.py:285-290)."""
    pos, types, box = _small_water(4, 9.0, seed=3)
    rc = 4.4
    for ej, ek in ((1, 1), (1, 2), (2, 2)):
        row = [3, ej, ek, 0.03, 0.0, 4.0, -1.0, rc, COS, 0]
        G = O.acsf(pos, types, box, [row])[:, 0]
        for i in range(len(types)):
            acc = 0.0
            nb = []
            for j in range(len(types)):
                d = pos[i] - pos[j]
                d -= box * np.round(d / box)
                r = np.linalg.norm(d)
                if 0 < r <= rc:
                    nb.append((j, d, r))
            for (a, (j, dj, rj)), (b, (k, dk, rk)) in itertools.combinations(enumerate(nb), 2):
                tj, tk = types[j], types[k]
                if {tj, tk} != {ej, ek} or (ej != ek and tj == tk):
                    continue
                djk = dj - dk
                djk -= box * np.round(djk / box)
                rjk = np.linalg.norm(djk)
                if rjk <= 0:
                    continue
                c = dj @ dk / (rj * rk)
                f = lambda r: 0.5 * (math.cos(math.pi * r / rc) + 1) if r < rc else 0.0
                acc += 2 ** (1 - 4.0) * (1 - c) ** 4 * math.exp(-0.03 * (rj * rj + rk * rk + rjk * rjk)) * f(rj) * f(rk) * f(rjk)
            assert G[i] == pytest.approx(acc, rel=1e-12, abs=1e-15)


def _g9_brute(pos, types, box, ej, ek, eta, zeta, lam, rc, fc):
    """G9 of every atom by an explicit loop over ORDERED pairs (j, k), k != j,
    following angular.py:92-107 and acsf.py:294-330: no cutoff on r_jk, no
    r_jk in the exponent, halved when the two neighbour elements coincide."""
    n = len(types)
    out = np.zeros(n)
    for i in range(n):
        nb = []
        for j in range(n):
            d = pos[i] - pos[j]
            if box is not None:
                d = d - box * np.round(d / box)
            r = np.linalg.norm(d)
            if 0 < r <= rc:
                nb.append((j, d, r))
        acc = 0.0
        for (j, dj, rj) in nb:
            if types[j] != ej:
                continue
            for (k, dk, rk) in nb:
                if types[k] != ek or k == j:
                    continue
                c = dj @ dk / (rj * rk)
                acc += 2.0 ** (1 - zeta) * (1 + lam * c) ** zeta * math.exp(-eta * (rj * rj + rk * rk)) * fc(rj) * fc(rk)
        out[i] = acc * (0.5 if ej == ek else 1.0)
    return out


@pytest.mark.parametrize("ej,ek,zeta,lam", [(1, 1, 4.0, -1.0), (1, 2, 1.0, 1.0), (2, 2, 2.0, 1.0), (1, 2, 3.0, -1.0)])
def test_g9_matches_ordered_pair_brute_force(ej, ek, zeta, lam):
    """G9 (angular.py:92-107) against an ordered-pair loop, with a cutoff
    larger than half the box so pairs with r_jk > rc occur (they count in G9
    and not in G3)."""
    pos, types, box = _small_water(4, 9.0, seed=3)
    rc = 5.5
    tanh_pre = O.tanh_prefactor()
    fc = lambda r: tanh_pre * math.tanh(1 - r / rc) ** 3 if r < rc else 0.0
    row = [9, ej, ek, 0.02, 0.0, zeta, lam, rc, TANH, 0]
    G = O.acsf(pos, types, box, [row])[:, 0]
    want = _g9_brute(pos, types, box, ej, ek, 0.02, zeta, lam, rc, fc)
    assert np.abs(want).max() > 1e-3
    assert np.allclose(G, want, rtol=1e-12, atol=1e-15)
    # a G9 that took fc(r_jk) or r_jk^2 into account, as G3 does, would differ
    g3 = O.acsf(pos, types, box, [[3] + row[1:]])[:, 0]
    assert np.abs(G - g3).max() > 1e-3 * np.abs(G).max()


def test_g9_three_atom_closed_form():
    """Centre O at the origin, H at (a, 0, 0) and O at (0, b, 0): cos = 0 and
    r_jk = sqrt(a^2 + b^2) > rc, so G3 = 0 (fc(r_jk) = 0) while
    G9 = 2^(1-zeta) exp(-eta (a^2 + b^2)) fc(a) fc(b) (distinct elements: no
    halving, one ordered pair); the same pair seen from the H centre has
    cos = a / r_jk and only one neighbour within rc (G9 = 0)."""
    a, b, rc, eta, zeta = 2.0, 3.0, 3.4, 0.07, 4.0
    pos = np.array([[0.0, 0.0, 0.0], [a, 0.0, 0.0], [0.0, b, 0.0]])
    types = np.array([2, 1, 2], dtype=np.int32)
    fc = lambda r: 0.5 * (math.cos(math.pi * r / rc) + 1) if r < rc else 0.0
    for lam in (1.0, -1.0):
        G9 = O.acsf(pos, types, None, [[9, 1, 2, eta, 0.0, zeta, lam, rc, COS, 0]])[:, 0]
        G3 = O.acsf(pos, types, None, [[3, 1, 2, eta, 0.0, zeta, lam, rc, COS, 0]])[:, 0]
        expect = 2.0 ** (1 - zeta) * math.exp(-eta * (a * a + b * b)) * fc(a) * fc(b)
        assert G9[0] == pytest.approx(expect, rel=1e-14)
        assert G9[1] == 0.0 and G9[2] == 0.0 and G3[0] == 0.0
    # inside the cutoff both see the pair; with cos = 0 the lambda sign drops out
    rc2 = 4.0
    fc2 = lambda r: 0.5 * (math.cos(math.pi * r / rc2) + 1)
    rjk = math.hypot(a, b)
    G3 = O.acsf(pos, types, None, [[3, 1, 2, eta, 0.0, 1.0, -1.0, rc2, COS, 0]])[:, 0]
    assert G3[0] == pytest.approx(math.exp(-eta * (a * a + b * b + rjk * rjk)) * fc2(a) * fc2(b) * fc2(rjk), rel=1e-14)


# ---------------------------------------------------------------------------
# finite differences and invariants
# ---------------------------------------------------------------------------

def test_selfjac_matches_fd_of_displaced_centre():
    pos, types, box = _small_water()
    rows = _teacher_rows()
    centers = np.arange(0, len(types), 5)
    J = O.acsf_selfjac(pos, types, box, rows, center_pos=pos[centers], self_idx=centers.astype(np.int32))
    h = 1e-5
    for d in range(3):
        e = np.zeros(3); e[d] = h
        gp = O.acsf(pos, types, box, rows, center_pos=pos[centers] + e, self_idx=centers.astype(np.int32))
        gm = O.acsf(pos, types, box, rows, center_pos=pos[centers] - e, self_idx=centers.astype(np.int32))
        fd = (gp - gm) / (2 * h)
        assert np.allclose(J[:, :, d], fd, rtol=1e-6, atol=1e-9)


def _random_nnp(rows_by_type, seed=0, hidden=6):
    rng = np.random.default_rng(seed)
    layers, scal = {}, {}
    for e, t in (("H", 1), ("O", 2)):
        n = rows_by_type[t].shape[0]
        layers[e] = [(rng.normal(size=(n, hidden)) / np.sqrt(n), np.zeros(hidden), "tanh"),
                     (rng.normal(size=(hidden, hidden)) / np.sqrt(hidden), np.zeros(hidden), "tanh"),
                     (rng.normal(size=(hidden, 1)) / np.sqrt(hidden), np.zeros(1), "identity")]
        scal[e] = {"mean": np.zeros(n), "sigma": np.ones(n), "minval": np.zeros(n), "maxval": np.ones(n) * 3}
    return O.OracleNNP(["H", "O"], {"H": 1, "O": 2}, {"H": rows_by_type[1], "O": rows_by_type[2]}, "scale", scal, layers)


def test_full_forces_are_minus_grad_energy():
    pos, types, box = _small_water()
    rows = _teacher_rows()
    nnp = _random_nnp({1: rows, 2: rows})
    E, _, f_self, f_full = nnp.energy_forces(pos, types, box)
    h = 1e-5
    for a in (0, 1, 7, 12, 23):
        for d in range(3):
            p = pos.copy(); p[a, d] += h
            m = pos.copy(); m[a, d] -= h
            fd = -(nnp.energy(p, types, box) - nnp.energy(m, types, box)) / (2 * h)
            assert f_full[a, d] == pytest.approx(fd, rel=1e-6, abs=1e-9)
    assert np.max(np.abs(f_full.sum(0))) < 1e-12 * np.abs(f_full).max() * len(types)


def test_self_forces_are_minus_grad_of_own_energy():
    """F_self,i = -dE_i/dr_i with neighbours frozen (force.py:16-43)."""
    pos, types, box = _small_water()
    rows = _teacher_rows()
    nnp = _random_nnp({1: rows, 2: rows}, seed=3)
    _, _, f_self, _ = nnp.energy_forces(pos, types, box)
    h = 1e-5
    for a in (0, 4, 11):
        e = "O" if types[a] == 2 else "H"
        def e_i(p_center):
            G = O.acsf(pos, types, box, nnp.tables[e], center_pos=p_center[None], self_idx=np.array([a], np.int32))
            x = O.scaler_transform("scale", nnp.scaler_params[e], G, 0.0, 1.0)
            return O.mlp_forward_backward(nnp.layers[e], x)[0][0]
        for d in range(3):
            dp = np.zeros(3); dp[d] = h
            fd = -(e_i(pos[a] + dp) - e_i(pos[a] - dp)) / (2 * h)
            assert f_self[a, d] == pytest.approx(fd, rel=1e-6, abs=1e-9)


def test_descriptor_rotation_invariance():
    pos, types, _ = _small_water(6, 16.0, seed=9)
    rows = _teacher_rows(rc=7.0)
    th = 0.7
    R = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1.0]])
    G0 = O.acsf(pos, types, None, rows)
    G1 = O.acsf(pos @ R.T + 3.0, types, None, rows)
    assert np.allclose(G0, G1, rtol=1e-12, atol=1e-14)


def test_reference_algorithm_matches_oracle():
    """The dense reference algorithm (timed as the CPU baseline) gives the oracle's
    descriptors and, with weights w, the self-term gradient sum_s w_s dG_is/dr_i."""
    pos, types, box = _small_water(6, 12.0, seed=4)
    rows = _teacher_rows(rc=5.0)
    centres = np.arange(0, len(types), 2).astype(np.int32)
    G, _ = O.reference_algorithm(pos, types, box, rows, centres)
    G0 = O.acsf(pos, types, box, rows, center_pos=pos[centres], self_idx=centres)
    np.testing.assert_allclose(G, G0, rtol=1e-12, atol=1e-14)
    rng = np.random.default_rng(2)
    w = rng.normal(size=G.shape)
    _, grad = O.reference_algorithm(pos, types, box, rows, centres, w)
    J = O.acsf_selfjac(pos, types, box, rows, center_pos=pos[centres], self_idx=centres)
    np.testing.assert_allclose(grad, np.einsum("cs,csd->cd", w, J), rtol=1e-10, atol=1e-12)


def test_cell_list_rows_and_centres_match_dense_oracle():
    """The cell-list CPU baseline: same rows as the dense mask, same descriptors and
    self gradients as or_acsf / the self-Jacobian."""
    pos, types, box = _small_water(40, 19.0, seed=8)
    rc = 6.0
    counts, idx = O.neighbors(pos, box, rc)
    row_off, idx2 = O.cell_neighbors(pos, box, rc)
    assert np.array_equal(np.diff(row_off), counts) and np.array_equal(idx2, idx)
    rows = _teacher_rows(rc=rc)
    centres = np.arange(1, len(types), 3).astype(np.int32)
    rng = np.random.default_rng(4)
    w = rng.normal(size=(centres.size, rows.shape[0]))
    G, g = O.celllist_centres(pos, types, box, row_off, idx2, rows, centres, w)
    np.testing.assert_allclose(G, O.acsf(pos, types, box, rows, center_pos=pos[centres], self_idx=centres),
                               rtol=1e-13, atol=1e-15)
    J = O.acsf_selfjac(pos, types, box, rows, center_pos=pos[centres], self_idx=centres)
    np.testing.assert_allclose(g, np.einsum("cs,csd->cd", w, J), rtol=1e-11, atol=1e-13)
