"""Training kernels (pt_train_seeds, pt_acsf_tangent, pt_mlp_train_grad,
pt_adam_step) against the oracle.

The loss gradient is pinned by central finite differences of the oracle's
FP64 loss  w_E((E-E*)/N)^2 + w_F/(3N) sum|F-F*|^2  (F = -grad E from
oracle.c) with respect to sampled weights and biases of every layer of every
element: an independent route that shares no adjoint code with the kernels.
"""

import math

import numpy as np
import pytest

from oracle.bridge import oracle_nnp

pytestmark = pytest.mark.gpu


def _setup(precision, n_mol=32, seed=5, hidden=(12, 10)):
    import torch
    from pantea_amd import water
    from pantea_amd.atoms import Structure
    from pantea_amd.bench_models import water_nnp
    from pantea_amd.student import StudentTrainer

    L = water.box_for_density(n_mol)
    s = Structure.from_dict(water.water_box(n_mol, L, seed=seed))
    pot = water_nnp(s, hidden=hidden, seed=seed, precision=precision, student=True)
    orc = oracle_nnp(pot, s)
    pos = s.positions.cpu().numpy()
    types = s.atom_types_host
    box = np.asarray(s.box_lengths, dtype=np.float64)
    E0, _, _, F0 = orc.energy_forces(pos, types, box)
    rng = np.random.default_rng(seed)
    # targets: a shifted energy and forces with a random perturbation so both terms are active
    E_ref = E0 + 0.05 * s.natoms
    F_ref = F0 + rng.normal(scale=np.std(F0) + 1e-3, size=F0.shape)
    tr = StudentTrainer(pot, precision=precision, w_energy=1.0, w_force=3.0)
    tr.add(s, E_ref, torch.as_tensor(F_ref, dtype=torch.float64, device="cuda"))
    tr.prepare(fit_scaler=False)
    return s, pot, orc, tr, (pos, types, box, E_ref, F_ref)


def _oracle_loss(orc, data, w_e, w_f):
    pos, types, box, E_ref, F_ref = data
    n = len(types)
    E, _, _, F = orc.energy_forces(pos, types, box)
    return w_e * ((E - E_ref) / n) ** 2, w_f * np.sum((F - F_ref) ** 2) / (3 * n)


def _fd_samples(orc, pot, rng, per_tensor=6):
    """(element, flat index, layer key, 'kernel'|'bias', position) for sampled parameters."""
    out = []
    for e in pot.elements:
        off = 0
        for li, (W, b, _a) in enumerate(orc.layers[e]):
            for name, arr in (("W", W), ("b", b)):
                k = min(per_tensor, arr.size)
                for flat in rng.choice(arr.size, size=k, replace=False):
                    out.append((e, off + int(flat), li, name, np.unravel_index(int(flat), arr.shape)))
                off += arr.size
    return out


def _fd(orc, data, w, sample, h):
    e, _, li, name, idx = sample
    W, b, a = orc.layers[e][li]
    arr = W if name == "W" else b
    keep = arr[idx]
    vals = []
    for sgn in (1.0, -1.0):
        arr[idx] = keep + sgn * h
        lE, lF = _oracle_loss(orc, data, *w)
        vals.append(lE + lF)
    arr[idx] = keep
    return (vals[0] - vals[1]) / (2 * h)


@pytest.mark.parametrize("precision,rtol", [("float64", 1e-6), ("mixed", 5e-3)])
def test_loss_gradient_matches_oracle_finite_differences(precision, rtol):
    s, pot, orc, tr, data = _setup(precision)
    loss, grads = tr.gradient([0])
    loss = loss.cpu().numpy()
    lE, lF = _oracle_loss(orc, data, tr.w_energy, tr.w_force)
    assert loss[0] == pytest.approx(lE, rel=1e-9 if precision == "float64" else 1e-5)
    assert loss[1] == pytest.approx(lF, rel=1e-9 if precision == "float64" else 1e-4)
    rng = np.random.default_rng(11)
    samples = _fd_samples(orc, pot, rng)
    g = {e: v.cpu().numpy() for e, v in grads.items()}
    scale = max(np.max(np.abs(v)) for v in g.values())
    errs = []
    for smp in samples:
        ref = _fd(orc, data, (tr.w_energy, tr.w_force), smp, h=1e-5)
        got = g[smp[0]][smp[1]]
        errs.append(abs(got - ref) / scale)
        assert abs(got - ref) <= rtol * scale, (smp, got, ref)
    print(f"{precision}: max |dL/dtheta - FD| / max|dL/dtheta| = {max(errs):.2e} over {len(samples)} parameters")


def test_loss_gradient_is_additive_over_structures():
    """The batch gradient is the mean of per-structure gradients (accumulation in pt_mlp_train_grad)."""
    import torch
    from pantea_amd import water
    from pantea_amd.atoms import Structure

    s, pot, orc, tr, data = _setup("float64")
    s2 = Structure.from_dict(water.water_box(40, water.box_for_density(40), seed=9))
    e2, f2 = pot.energy_and_forces(s2, mode="full")
    tr.add(s2, float(e2) - 1.0, f2 * 0.5)
    tr.prepare(fit_scaler=False)
    l0, g0 = tr.gradient([0])
    l1, g1 = tr.gradient([1])
    l01, g01 = tr.gradient([0, 1])
    torch.testing.assert_close(l01, 0.5 * (l0 + l1), rtol=1e-12, atol=0)
    for e in g0:
        torch.testing.assert_close(g01[e], 0.5 * (g0[e] + g1[e]), rtol=1e-10, atol=1e-14)


def test_adam_matches_closed_form():
    import torch
    from pantea_amd import _lib

    dev = _lib.require_device()
    rng = np.random.default_rng(3)
    n = 1000
    p = rng.normal(size=n)
    m = np.zeros(n)
    v = np.zeros(n)
    P = torch.as_tensor(p, device=dev)
    M = torch.zeros(n, dtype=torch.float64, device=dev)
    V = torch.zeros(n, dtype=torch.float64, device=dev)
    lr, b1, b2, eps, wd = 1e-2, 0.9, 0.999, 1e-8, 1e-3
    for t in range(1, 6):
        g = rng.normal(size=n)
        G = torch.as_tensor(g, device=dev)
        _lib.check(_lib.load().pt_adam_step(n, P.data_ptr(), G.data_ptr(), M.data_ptr(), V.data_ptr(), lr, b1,
                                            b2, eps, wd, t, _lib.stream_ptr()), "adam")
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        mh = m / (1 - b1 ** t)
        vh = v / (1 - b2 ** t)
        p = p - lr * (mh / (np.sqrt(vh) + eps) + wd * p)
    np.testing.assert_allclose(P.cpu().numpy(), p, rtol=1e-13, atol=1e-15)


def test_training_reduces_loss_and_force_error():
    """A few hundred Adam steps on one structure lower the force RMSE against fixed targets."""
    s, pot, orc, tr, data = _setup("mixed", hidden=(16, 16))
    before = tr.evaluate([0])
    tr.lr = 3e-3
    for _ in range(200):
        tr.step([0])
    tr.sync_params()
    after = tr.evaluate([0])
    assert after["force_rmse_mev_per_bohr"] < 0.8 * before["force_rmse_mev_per_bohr"], (before, after)
    # the trained parameters round-trip into the potential's own evaluation path
    e, f = pot.energy_and_forces(s, mode="full", precision="mixed")
    import torch
    F_ref = torch.as_tensor(data[4], device=f.device)
    rmse = math.sqrt(float(((f - F_ref) ** 2).mean()))
    from pantea_amd.units import units
    assert rmse * units.TO_MILLI_ELECTRON_VOLT == pytest.approx(after["force_rmse_mev_per_bohr"], rel=1e-6)


def test_trained_student_gains_over_untrained_network():
    """What the learned part buys: with a teacher the student can represent (the
    student's descriptors and layer widths, other weights), GPU training on the
    teacher's own MD data cuts the student's force error several-fold, and in
    surrogate MD the trained student passes the judge and drives the run, while
    the same network untrained is rejected at every judge and the teacher drives
    everything."""
    import torch
    from pantea_amd import water
    from pantea_amd.atoms import Structure
    from pantea_amd.bench_models import make_nnp, student_acsf
    from pantea_amd.simulation import BrendsenThermostat
    from pantea_amd.student import StudentTrainer, teacher_snapshots
    from pantea_amd.surrogate import SurrogateMD
    from pantea_amd.units import units

    mev = units.TO_MILLI_ELECTRON_VOLT
    s0 = Structure.from_dict(water.water_box(64, water.box_for_density(64), seed=21))
    teacher = make_nnp(student_acsf, (16, 16), seed=11, precision="float64")
    teacher.fit_scaler([s0])
    # a smoother teacher (kernels x 0.5) whose scaler the student shares
    teacher.set_model_params({e: {k: {"kernel": 0.5 * d["kernel"], "bias": d["bias"]} for k, d in prm.items()}
                              for e, prm in teacher.models_params.items()})
    student = make_nnp(student_acsf, (16, 16), seed=12, precision="mixed")
    student.scalers_params = dict(teacher.scalers_params)
    untrained_params = {e: {k: {p: v.copy() for p, v in d.items()} for k, d in prm.items()}
                        for e, prm in student.models_params.items()}
    tr = StudentTrainer(student, precision="mixed", w_energy=1.0, w_force=10.0, lr=5e-3, seed=1)
    # 4 teacher trajectories of 120 steps (300 K), a snapshot every 5 steps
    for snap, e, f in teacher_snapshots(teacher, n_structures=4, n_snapshots=96, every=5, min_molecules=48,
                                        max_molecules=72, seed=3):
        tr.add(snap, e, f)
    tr.prepare(fit_scaler=False)
    idx = np.arange(len(tr.snapshots))
    val, train = idx[::6], np.setdiff1d(idx, idx[::6])
    before = tr.evaluate(val)["force_rmse_mev_per_bohr"]
    tr.train(600, batch=8, train_idx=train)
    after = tr.evaluate(val)["force_rmse_mev_per_bohr"]
    print(f"validation force RMSE: untrained {before:.1f}, trained {after:.1f} meV/Bohr")
    assert after < 0.4 * before

    untrained = make_nnp(student_acsf, (16, 16), seed=12, precision="mixed")
    untrained.set_model_params(untrained_params)
    untrained.scalers_params = dict(student.scalers_params)
    thr = math.sqrt(before * after) / mev
    runs = {}
    for name, stu in (("trained", student), ("untrained", untrained)):
        sur = SurrogateMD(teacher, stu, s0, seed=5, judge_every=10, threshold=thr, fallback_steps=20, skin=1.0,
                          thermostat=BrendsenThermostat(300.0, 0.5 * units.FROM_FEMTO_SECOND))
        runs[name] = sur.run(120).as_dict()
        print(name, runs[name])
    assert runs["untrained"]["failures"] == runs["untrained"]["judges"] > 0
    assert runs["trained"]["teacher_share"] < 0.5 * runs["untrained"]["teacher_share"]
    assert runs["trained"]["student_steps"] - runs["trained"]["discarded_steps"] > 60
