"""The learned part at the benchmark configurations.

* Student training on teacher data of the student-training configuration's
  shape (water boxes of U{32..512} molecules at 1 g/cm^3, a snapshot every
  10 steps of the 98,304-atom teacher's 300 K MD, FP64 labels), at a reduced
  snapshot count: the loss gradient on the largest structure against central
  finite differences of the oracle's FP64 loss along random directions in
  parameter space, and the training loss falls.
* Surrogate MD at 98,304 atoms with the packaged student and the benchmark
  teacher for 5,000 steps (judge every 20 steps on 128 sampled atoms,
  200-step fallback, 5 meV/Bohr): the bookkeeping invariants, and every kept
  step lies in a window whose judge RMSE is within the threshold (or was
  driven by the judge's own model).
* The same surrogate with the threshold at the student's own accuracy
  (twice its force RMSE on the start structure): the trained student passes
  judges and drives, the untrained network with the same descriptors does not.
"""

import math

import numpy as np
import pytest

from oracle.bridge import oracle_nnp

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def teacher98k():
    from pantea_amd import bench_models

    return bench_models.water_teacher(precision="mixed")


def _oracle_loss(orc, pos, types, box, E_ref, F_ref, w_e, w_f):
    n = len(types)
    E, _, _, F = orc.energy_forces(pos, types, box)
    return w_e * ((E - E_ref) / n) ** 2 + w_f * np.sum((F - F_ref) ** 2) / (3 * n)


@pytest.mark.parametrize("precision,rtol", [("float64", 1e-6), ("mixed", 5e-3)])
def test_student_training_at_the_training_configuration(teacher98k, precision, rtol):
    from pantea_amd import bench_models
    from pantea_amd.student import StudentTrainer, teacher_snapshots

    _, teacher = teacher98k
    data = list(teacher_snapshots(teacher, n_structures=8, n_snapshots=24, every=10, min_molecules=32,
                                  max_molecules=512, seed=77))
    sizes = sorted({s.natoms for s, _, _ in data})
    print(f"{len(data)} snapshots of {sizes} atoms", flush=True)
    assert len(data) == 24 and sizes[0] >= 96 and sizes[-1] <= 1536 and len(sizes) >= 4
    student = bench_models.water_student(None, seed=12, precision=precision)
    tr = StudentTrainer(student, precision=precision, w_energy=1.0, w_force=10.0, lr=1e-3, seed=3)
    for snap, e, f in data:
        tr.add(snap, e, f)
    tr.prepare(fit_scaler=True)
    big = int(np.argmax([sn.n for sn in tr.snapshots]))
    loss, grads = tr.gradient([big])
    g = {e: v.cpu().numpy() for e, v in grads.items()}
    # the oracle with the student's parameters, on the largest structure
    s_big, E_ref, F_ref = data[big]
    orc = oracle_nnp(student, s_big)
    pos = s_big.positions.cpu().numpy()
    types = s_big.atom_types_host
    box = np.asarray(s_big.box_lengths)
    F_ref = F_ref.cpu().numpy()
    L0 = _oracle_loss(orc, pos, types, box, E_ref, F_ref, tr.w_energy, tr.w_force)
    assert float(loss.sum()) == pytest.approx(L0, rel=1e-9 if precision == "float64" else 1e-4)
    # directional derivatives: two random directions and the gradient's own
    rng = np.random.default_rng(5)
    flat = {e: np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b, _ in orc.layers[e]])
            for e in orc.elements}
    dirs = [{e: rng.normal(size=v.size) for e, v in flat.items()} for _ in range(2)]
    dirs.append({e: g[e] for e in flat})
    h = 1e-5
    for u in dirs:
        norm = math.sqrt(sum(float(np.sum(v * v)) for v in u.values()))
        u = {e: v / norm for e, v in u.items()}
        vals = []
        for sgn in (1.0, -1.0):
            for e in orc.elements:
                o = 0
                for li, (W, b, _) in enumerate(orc.layers[e]):
                    W += sgn * h * u[e][o:o + W.size].reshape(W.shape)
                    o += W.size
                    b += sgn * h * u[e][o:o + b.size]
                    o += b.size
            vals.append(_oracle_loss(orc, pos, types, box, E_ref, F_ref, tr.w_energy, tr.w_force))
            for e in orc.elements:   # back to the unperturbed parameters
                o = 0
                for li, (W, b, _) in enumerate(orc.layers[e]):
                    W -= sgn * h * u[e][o:o + W.size].reshape(W.shape)
                    o += W.size
                    b -= sgn * h * u[e][o:o + b.size]
                    o += b.size
        fd = (vals[0] - vals[1]) / (2 * h)
        got = sum(float(np.dot(g[e], u[e])) for e in u)
        gnorm = math.sqrt(sum(float(np.sum(v * v)) for v in g.values()))
        print(f"{precision}: directional derivative {got:.10e} vs FD {fd:.10e}", flush=True)
        assert abs(got - fd) <= rtol * gnorm
    # and training lowers the loss
    hist = tr.train(8, batch=8)
    first, last = hist[0][1] + hist[0][2], hist[-1][1] + hist[-1][2]
    print(f"{precision}: loss {first:.4e} -> {last:.4e} over 8 epochs", flush=True)
    assert last < 0.7 * first


def _surrogate(teacher, student, s, threshold_mev, steps=200, judge_atoms=None):
    from pantea_amd.simulation import BrendsenThermostat
    from pantea_amd.surrogate import SurrogateMD
    from pantea_amd.units import units

    dt = 0.5 * units.FROM_FEMTO_SECOND
    sur = SurrogateMD(teacher, student, s, seed=2023, judge_every=20, fallback_steps=200,
                      threshold=threshold_mev / units.TO_MILLI_ELECTRON_VOLT, fallback_precision="mixed",
                      skin=1.0, thermostat=BrendsenThermostat(300.0, dt), judge_atoms=judge_atoms,
                      probation_every=5, graph=True)
    st = sur.run(steps, judge_tail=True)
    return sur, st


def _check_windows(st, steps, threshold_mev):
    from pantea_amd.units import units

    thr = threshold_mev / units.TO_MILLI_ELECTRON_VOLT
    w = sorted(st.windows)
    assert st.steps == steps
    assert w and w[0][0] == 0 and w[-1][1] == steps
    for (a0, b0, _, _), (a1, b1, _, _) in zip(w, w[1:]):
        assert b0 == a1
    for a, b, driver, rmse in w:
        assert b > a
        assert driver in ("student", "fallback", "reference")
        if driver == "reference":
            assert rmse is None
        else:
            assert rmse <= thr
    assert st.judges == len(st.judge_rmse) == len(st.judge_driver)
    assert st.failures == sum(1 for r in st.judge_rmse if not r <= thr)
    assert st.fallback_steps + sum(b - a for a, b, d, _ in w if d == "student") == steps


def test_surrogate_98k_at_the_configured_threshold(teacher98k):
    from pantea_amd import bench_models
    from pantea_amd.student import ASSET_DIR

    s, teacher = teacher98k
    student = bench_models.water_student(ASSET_DIR)
    sur, st = _surrogate(teacher, student, s, 5.0, steps=5000, judge_atoms=128)
    print(st.as_dict(), flush=True)
    _check_windows(st, 5000, 5.0)


def test_trained_student_drives_at_its_own_accuracy(teacher98k):
    from pantea_amd import bench_models
    from pantea_amd.student import ASSET_DIR
    from pantea_amd.units import units

    s, teacher = teacher98k
    student = bench_models.water_student(ASSET_DIR)
    e64, f64 = teacher.energy_and_forces(s, mode="full", precision="float64")
    _, fs = student.energy_and_forces(s, mode="full", precision="mixed")
    rmse = float(torch.sqrt(torch.mean((fs - f64) ** 2))) * units.TO_MILLI_ELECTRON_VOLT
    thr = 2.0 * rmse
    sur, st = _surrogate(teacher, student, s, thr, judge_atoms=128)
    d = st.as_dict()
    print(f"threshold {thr:.1f} meV/Bohr: {d}", flush=True)
    _check_windows(st, 200, thr)
    assert d["student_kept_steps"] > 0
    # the same descriptors and widths, untrained: rejected at every judge
    untrained = bench_models.student_from_spec(student.student_spec, seed=12)
    untrained.scalers_params = dict(student.scalers_params)
    sur_u, st_u = _surrogate(teacher, untrained, s, thr, judge_atoms=128)
    du = st_u.as_dict()
    print(f"untrained: {du}", flush=True)
    assert du["student_kept_steps"] < d["student_kept_steps"]
