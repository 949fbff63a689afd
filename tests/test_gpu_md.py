"""Velocity Verlet, thermostat and surrogate-driven MD on the GPU against the oracle.

One MD step is checked against the oracle's closed forms
(molecular_dynamics.py:16-30, thermostat.py:12-22) fed with the oracle's own
FP64 forces; the long runs check invariants (NVE energy drift, the surrogate
driver's bookkeeping and that a rewound, teacher-driven run is the teacher's
trajectory).
"""

import math

import numpy as np
import pytest

from oracle import pyoracle
from oracle.bridge import oracle_nnp

pytestmark = pytest.mark.gpu


def _cfg0(hidden=(25, 25), n_mol=128, L=29.56, seed=2023):
    from pantea_amd import water
    from pantea_amd.atoms import Structure
    from pantea_amd.bench_models import water_nnp

    s = Structure.from_dict(water.water_box(n_mol, L, seed=seed))
    return s, water_nnp(s, hidden=hidden, seed=seed, precision="float64")


@pytest.mark.parametrize("integrator,force_mode", [("physical", "full"), ("reference", "reference")])
@pytest.mark.parametrize("skin", [0.0, 1.0])
def test_verlet_step_matches_oracle(integrator, force_mode, skin):
    from pantea_amd.simulation import BrendsenThermostat, MDSimulator, System
    from pantea_amd.units import units

    s, pot = _cfg0()
    orc = oracle_nnp(pot, s)
    sysm = System.from_structure(s, pot, temperature=300.0, seed=11, force_mode=force_mode,
                                 precision="float64", skin=skin)
    dt = 0.5 * units.FROM_FEMTO_SECOND
    md = MDSimulator(dt, integrator=integrator)
    x0 = sysm.positions.cpu().numpy().copy()
    v0 = sysm.velocities.cpu().numpy().copy()
    m = sysm.masses.cpu().numpy().reshape(-1, 1)
    types = s.atom_types_host
    box = np.asarray(s.box_lengths)
    col = 3 if force_mode == "full" else 2

    def acc(x):
        f = orc.energy_forces(x, types, box)[col]
        return f / m if integrator == "physical" else f, f

    # each step is checked as a map from the GPU's state before it, so that rounding does not
    # compound through the trajectory
    a0, _ = acc(x0)
    for step in range(3):
        md.simulate_one_step(sysm)
        x1 = pyoracle.verlet_positions(x0, v0, a0, dt, box)
        xg = sysm.positions.cpu().numpy()
        d = xg - x1
        d -= box * np.round(d / box)          # a coordinate may wrap on the other side of 0 / L
        # the reference integrator (a = F, no mass) moves atoms by tens of Bohr per step, so the
        # forces' 1e-10-level rounding shows up scaled by the step length
        disp = np.max(np.abs(v0 * dt + 0.5 * a0 * dt * dt))
        assert np.max(np.abs(d)) < 1e-11 + 1e-9 * disp, (step, np.max(np.abs(d)), disp)
        a1, f1 = acc(xg)
        v1 = pyoracle.verlet_velocities(v0, a0, a1, dt)
        # componentwise relative to the largest component: cancelling components carry the
        # rounding of their large terms
        np.testing.assert_allclose(sysm.forces.cpu().numpy(), f1, rtol=0, atol=1e-10 * np.max(np.abs(f1)))
        np.testing.assert_allclose(sysm.velocities.cpu().numpy(), v1, rtol=0, atol=1e-10 * np.max(np.abs(v1)))
        x0, v0, a0 = xg.copy(), sysm.velocities.cpu().numpy().copy(), a1
    assert md.step == 3 and md.elapsed_time == pytest.approx(3 * dt)


@pytest.mark.parametrize("precision,tol", [("float64", 1e-10), ("mixed", 3e-3)])
def test_md_forces_with_distinct_element_tables(precision, tol):
    """The MD engine launches the descriptor kernels once for elements whose tables are
    byte-identical (the teacher's O and H) and once per table otherwise: with an O table
    that differs from the H table (other radial eta values, other function counts) its
    -grad E matches the oracle's."""
    from pantea_amd.bench_models import make_nnp
    from pantea_amd.descriptors import ACSF, G2, G3, CutoffFunction, NeighborElements
    from pantea_amd.simulation import System

    s, _ = _cfg0()
    cfn = CutoffFunction.from_type("tanh", 9.0)

    def acsf_of(el):
        etas = (0.02, 0.05) if el == "O" else (0.01, 0.03, 0.08)
        radial = tuple((G2(cfn, 0.0, eta), NeighborElements(e)) for e in ("H", "O") for eta in etas)
        angular = tuple((G3(cfn, eta=0.005, zeta=z, lambda0=lam, r_shift=0.0), NeighborElements(j, k))
                        for (j, k) in (("H", "H"), ("H", "O"), ("O", "O")) for z in (1.0, 4.0)
                        for lam in (1.0, -1.0))
        return ACSF(el, radial, angular)

    pot = make_nnp(acsf_of, (8, 8), seed=5, precision="float64")
    pot.fit_scaler([s])
    _, _, _, f_full = oracle_nnp(pot, s).energy_forces(s.positions.cpu().numpy(), s.atom_types_host,
                                                       s.box_lengths)
    sysm = System.from_structure(s, pot, temperature=300.0, seed=3, force_mode="full", precision=precision,
                                 skin=1.0)
    eng = sysm._engine()
    assert len(eng.ff._desc_runs(eng.nl)) == 2
    sysm.update_forces_from_positions()
    f = sysm.structure.forces.cpu().numpy()
    err = float(np.max(np.abs(f - f_full)) / np.max(np.abs(f_full)))
    print(f"{precision}: max force error {err:.2e} of max |F|")
    assert err < tol


@pytest.mark.parametrize("n", [1, 2, 300, 100003])
def test_max_displacement_top_two(n):
    """pt_max_displacement: the largest and second largest minimum-image displacement
    (two different atoms), which bound every pair's relative motion for the skin-list
    rebuild test; displacements up to 0.6 box lengths so the image shift acts."""
    import ctypes

    import torch

    from pantea_amd import _lib
    from pantea_amd.atoms import host_box

    rng = np.random.default_rng(n)
    L = np.array([20.0, 23.0, 27.0])
    ref = rng.uniform(0.0, 1.0, (n, 3)) * L
    d = rng.normal(0.0, 0.3, (n, 3))
    d[rng.integers(0, n)] += 0.6 * L   # one far move, wrapped by the image shift
    pos = ref + d
    pos4 = np.zeros((n, 4)); pos4[:, :3] = pos
    ref4 = np.zeros((n, 4)); ref4[:, :3] = ref
    out = torch.zeros(514, dtype=torch.float64, device="cuda")
    box = host_box(L)
    p4, r4 = torch.tensor(pos4, device="cuda"), torch.tensor(ref4, device="cuda")
    lib = _lib.load()
    _lib.check(lib.pt_max_displacement(n, p4.data_ptr(), r4.data_ptr(), ctypes.cast(box, ctypes.c_void_p),
                                       out.data_ptr(), _lib.stream_ptr()), "pt_max_displacement")
    got = out[:2].cpu().numpy()
    dd = pos - ref
    dd = np.where(dd > 0.5 * L, dd - L, dd)
    dd = np.where(dd < -0.5 * L, dd + L, dd)
    r = np.sort(np.sqrt((dd * dd).sum(1)))[::-1]
    want = np.array([r[0], r[1] if n > 1 else 0.0])
    assert np.allclose(got, want, rtol=1e-14, atol=0.0), (got, want)


def test_berendsen_step_matches_oracle():
    from pantea_amd.simulation import BrendsenThermostat, MDSimulator, System
    from pantea_amd.units import units

    s, pot = _cfg0()
    orc = oracle_nnp(pot, s)
    sysm = System.from_structure(s, pot, temperature=450.0, seed=3, force_mode="full", precision="float64")
    dt = 0.5 * units.FROM_FEMTO_SECOND
    tau = 10.0 * dt
    md = MDSimulator(dt, thermostat=BrendsenThermostat(300.0, tau), integrator="physical")
    x0 = sysm.positions.cpu().numpy().copy()
    v0 = sysm.velocities.cpu().numpy().copy()
    m = sysm.masses.cpu().numpy().reshape(-1, 1)
    types, box = s.atom_types_host, np.asarray(s.box_lengths)
    a0 = orc.energy_forces(x0, types, box)[3] / m
    md.simulate_one_step(sysm)
    x1 = pyoracle.verlet_positions(x0, v0, a0, dt, box)
    a1 = orc.energy_forces(x1, types, box)[3] / m
    v1 = pyoracle.verlet_velocities(v0, a0, a1, dt)
    lam = pyoracle.berendsen_factor(dt, tau, pyoracle.temperature(v1, m), 300.0)
    assert lam < 1.0
    np.testing.assert_allclose(sysm.velocities.cpu().numpy(), lam * v1, rtol=1e-9, atol=1e-14)
    assert float(sysm.get_temperature()) == pytest.approx(pyoracle.temperature(lam * v1, m), rel=1e-10)


def test_maxwell_boltzmann_velocities():
    """T = 300 K exactly before the COM shift (system.py:83-96), zero COM velocity after it."""
    from pantea_amd.simulation import System

    s, pot = _cfg0()
    sysm = System.from_structure(s, pot, temperature=300.0, seed=5, force_mode="full")
    v = sysm.velocities.cpu().numpy()
    m = sysm.masses.cpu().numpy().reshape(-1, 1)
    np.testing.assert_allclose((m * v).sum(0) / m.sum(), 0.0, atol=1e-15)
    assert float(sysm.get_temperature()) == pytest.approx(300.0, rel=0.02)


def test_nve_energy_error_is_second_order_in_dt():
    """Velocity Verlet with F = -grad E: the total-energy error after 5 fs falls as dt^2.

    The random-weight water NNP with its min-max scaler heats the box from 300 K
    to ~5e4 K within those 5 fs, which makes this a demanding check of the
    integrator and of F = -grad E together.
    """
    from pantea_amd.simulation import MDSimulator, System
    from pantea_amd.units import units

    s, pot = _cfg0()
    errs = []
    for dt_fs in (0.5, 0.25, 0.125):
        sysm = System.from_structure(s, pot, temperature=300.0, seed=2023, force_mode="full",
                                     precision="float64", skin=1.0)
        md = MDSimulator(dt_fs * units.FROM_FEMTO_SECOND, integrator="physical")
        e0 = float(sysm.get_total_energy())
        md.run(sysm, int(round(5.0 / dt_fs)))
        errs.append(abs(float(sysm.get_total_energy()) - e0))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    print(f"|dE| at 5 fs for dt 0.5/0.25/0.125 fs: {errs}; ratios {r1:.2f} {r2:.2f}")
    assert 3.0 < r1 < 5.5 and 3.0 < r2 < 5.5


def _nve_drift(s, pot, precision, dt_fs, steps, chunks=10):
    """NVE run: per chunk, E(t) - E(0) per atom (meV) and T (K)."""
    from pantea_amd.simulation import MDSimulator, System
    from pantea_amd.units import units

    sysm = System.from_structure(s, pot, temperature=300.0, seed=2023, force_mode="full", precision=precision,
                                 skin=1.0, graph=True)
    md = MDSimulator(dt_fs * units.FROM_FEMTO_SECOND, integrator="physical")
    e0 = float(sysm.get_total_energy())
    n = s.natoms
    mev = units.TO_MILLI_ELECTRON_VOLT
    trace = []
    for k in range(chunks):
        md.run(sysm, steps // chunks)
        e = float(sysm.get_total_energy())
        assert math.isfinite(e)
        trace.append(((e - e0) / n * mev, float(sysm.get_temperature())))
    assert md.step == steps
    return trace, sysm


def _check_nve(name, runs, steps):
    """The drift stays a small fraction of the kinetic energy the random-weight NNP releases, and
    mixed precision follows FP64 while the two trajectories still coincide (first chunk)."""
    from pantea_amd.units import units

    kb_mev = 3.166811563e-6 * units.TO_MILLI_ELECTRON_VOLT   # Boltzmann constant in meV/K
    for prec, trace in runs.items():
        worst = max(abs(d) for d, _ in trace)
        ke = 1.5 * kb_mev * max(t for _, t in trace)          # kinetic energy per atom at the hottest point
        print(f"NVE {name} {prec}: max |dE| {worst:.4g} meV/atom = {worst / ke:.2e} of the kinetic energy "
              f"({ke:.4g} meV/atom); " + "; ".join(f"{(k + 1) * steps // len(trace)}: {d:.4g} meV/atom "
                                                     f"{t:.0f} K" for k, (d, t) in enumerate(trace)))
        assert worst <= 0.05 * ke
    d64, dm = runs["float64"][0][0], runs["mixed"][0][0]
    assert abs(dm - d64) <= 0.02 * max(abs(d64), 1.0)


def test_nve_cfg1_1000_steps_bounded_drift():
    """1000 NVE steps (a = F/m, F = -grad E, dt 0.0625 fs) on 4096 H2O in FP64 and mixed precision.
    The random-weight NNP turns ~6 eV/atom of potential energy into heat within 20 fs (300 K ->
    ~5e4 K, then ~9e4 K); the integrator's error is second order in dt (see above) and stays
    within 5 % of the kinetic energy over the whole run."""
    from pantea_amd import water
    from pantea_amd.atoms import Structure
    from pantea_amd.bench_models import water_nnp

    s = Structure.from_dict(water.water_box(4096, 93.85, seed=2023))
    runs = {}
    for prec in ("float64", "mixed"):
        pot = water_nnp(s, hidden=(25, 25), seed=2023, precision=prec)
        runs[prec], sysm = _nve_drift(s, pot, prec, 0.0625, 1000)
        assert sysm._engine().n_builds > 1
    _check_nve("cfg1", runs, 1000)


def test_nve_cfg2_500_steps_bounded_drift():
    """500 NVE steps (dt 0.0625 fs) of the 98,304-atom benchmark system and teacher, FP64 and mixed."""
    from pantea_amd import bench_models

    runs = {}
    s, pot = bench_models.water_teacher(precision="float64")
    for prec in ("float64", "mixed"):
        runs[prec], _ = _nve_drift(s, pot, prec, 0.0625, 500, chunks=5)
    _check_nve("cfg2", runs, 500)


def test_rewind_restores_state_exactly():
    from pantea_amd.simulation import MDSimulator, System
    from pantea_amd.units import units

    s, pot = _cfg0()
    sysm = System.from_structure(s, pot, temperature=300.0, seed=2, force_mode="full", precision="float64",
                                 skin=1.0)
    md = MDSimulator(0.5 * units.FROM_FEMTO_SECOND, integrator="physical")
    md.run(sysm, 5)
    eng = sysm._engine()
    st = eng.save_state()
    x_saved = sysm.positions.clone()
    md.run(sysm, 7)
    x_a = sysm.positions.clone()
    eng.restore_state(st)
    assert float((sysm.positions - x_saved).abs().max()) == 0.0
    md.run(sysm, 7)
    assert float((sysm.positions - x_a).abs().max()) < 1e-12


def test_surrogate_with_teacher_as_student_never_falls_back():
    from pantea_amd.simulation import MDSimulator, System
    from pantea_amd.surrogate import SurrogateMD
    from pantea_amd.units import units

    s, pot = _cfg0()
    sur = SurrogateMD(pot, pot, s, seed=4, student_precision="float64", judge_every=5, skin=1.0)
    st = sur.run(30)
    assert st.steps == 30 and st.judges == 6 and st.failures == 0 and st.fallback_steps == 0
    assert max(st.judge_rmse) < 1e-12
    plain = System.from_structure(s, pot, temperature=300.0, seed=4, force_mode="full", precision="float64",
                                  skin=1.0)
    MDSimulator(0.5 * units.FROM_FEMTO_SECOND, integrator="physical").run(plain, 30)
    assert float((sur.positions - plain.positions).abs().max()) < 1e-10


def test_surrogate_failed_judge_rewinds_to_teacher_trajectory():
    """threshold 0: every student window is rejected, so the kept trajectory is the teacher's own."""
    from pantea_amd.bench_models import make_nnp, student_acsf
    from pantea_amd.simulation import MDSimulator, System
    from pantea_amd.surrogate import SurrogateMD
    from pantea_amd.units import units

    s, teacher = _cfg0()
    student = make_nnp(student_acsf, (16, 16), seed=99, precision="mixed")
    student.fit_scaler([s])
    sur = SurrogateMD(teacher, student, s, seed=8, judge_every=4, threshold=0.0, fallback_steps=10, skin=1.0)
    st = sur.run(25)
    # windows: 4 student steps rejected, 10 teacher steps, ... -> kept steps all teacher-driven
    assert st.steps == 25
    assert st.fallback_steps == 25
    assert st.failures == st.judges == 3
    assert st.discarded_steps == 12
    plain = System.from_structure(s, teacher, temperature=300.0, seed=8, force_mode="full", precision="float64",
                                  skin=1.0)
    MDSimulator(0.5 * units.FROM_FEMTO_SECOND, integrator="physical").run(plain, 25)
    assert float((sur.positions - plain.positions).abs().max()) < 1e-9


def test_surrogate_probation_judges_a_failed_student_sooner():
    """A student is judged every probation_every steps until it passes a judge (from the
    start, and again after each failure): threshold 0 rejects every window after 2 student
    steps, where judge_every would take 4."""
    from pantea_amd.bench_models import make_nnp, student_acsf
    from pantea_amd.simulation import MDSimulator, System
    from pantea_amd.surrogate import SurrogateMD
    from pantea_amd.units import units

    s, teacher = _cfg0()
    student = make_nnp(student_acsf, (16, 16), seed=99, precision="mixed")
    student.fit_scaler([s])
    sur = SurrogateMD(teacher, student, s, seed=8, judge_every=4, threshold=0.0, fallback_steps=10, skin=1.0,
                      probation_every=2)
    st = sur.run(25)
    # 2 student steps rejected, 10 teacher, 2 rejected, 10 teacher, 2 rejected, 5 teacher
    assert st.steps == 25 and st.fallback_steps == 25
    assert st.failures == st.judges == 3
    assert st.discarded_steps == 2 + 2 + 2
    plain = System.from_structure(s, teacher, temperature=300.0, seed=8, force_mode="full", precision="float64",
                                  skin=1.0)
    MDSimulator(0.5 * units.FROM_FEMTO_SECOND, integrator="physical").run(plain, 25)
    assert float((sur.positions - plain.positions).abs().max()) < 1e-9


def test_surrogate_mixed_fallback_is_judged_and_drives():
    """A mixed-precision fallback under the FP64 judge: the untrained student is rejected at
    every judge, the mixed teacher passes its own judges and drives the kept trajectory,
    which is then the plain mixed-teacher MD; the FP64 model never drives."""
    from pantea_amd.bench_models import make_nnp, student_acsf
    from pantea_amd.simulation import MDSimulator, System
    from pantea_amd.surrogate import SurrogateMD
    from pantea_amd.units import units

    from pantea_amd.simulation import BrendsenThermostat

    s, teacher = _cfg0()
    student = make_nnp(student_acsf, (16, 16), seed=99, precision="mixed")
    student.fit_scaler([s])
    # 5 meV/Bohr: the mixed teacher stays within ~1 of FP64 at 300 K, the untrained
    # student is off by ~1e4
    thr = 5.0 / units.TO_MILLI_ELECTRON_VOLT
    thermo = BrendsenThermostat(300.0, 0.5 * units.FROM_FEMTO_SECOND)
    sur = SurrogateMD(teacher, student, s, seed=8, judge_every=4, threshold=thr, fallback_steps=10,
                      fallback_precision="mixed", skin=1.0, thermostat=thermo)
    st = sur.run(25)
    print("judge RMSE (meV/Bohr):", [round(r * units.TO_MILLI_ELECTRON_VOLT, 3) for r in st.judge_rmse],
          st.judge_driver)
    assert st.steps == 25 and st.fallback_steps == 25 and st.reference_steps == 0
    assert st.fallback_failures == 0
    # student rejected at steps 4, 14, 24; the mixed windows judged at 4, 8, 10 / 14, 18, 20 / 24
    assert st.failures == 3 and st.discarded_steps == 12
    assert st.judges == 10
    plain = System.from_structure(s, teacher, temperature=300.0, seed=8, force_mode="full", precision="mixed",
                                  skin=1.0)
    MDSimulator(0.5 * units.FROM_FEMTO_SECOND, thermostat=thermo, integrator="physical").run(plain, 25)
    # The rewinds rebuild the neighbour list at other steps than the plain run, which
    # reorders atoms and so the FP32 sums: the two mixed trajectories agree to FP32
    # rounding grown over 25 steps (~1e-5 Bohr), not bit for bit. A student-driven
    # step would move atoms by orders of magnitude more.
    assert float((sur.positions - plain.positions).abs().max()) < 1e-3


@pytest.mark.parametrize("student_rc", [12.0, 8.0])
def test_surrogate_failed_mixed_fallback_escalates_to_judge_model(student_rc):
    """threshold 0: the mixed fallback fails its first judge too, the run rewinds and the FP64
    model drives; the kept trajectory is the FP64 teacher's.  With an 8 Bohr student the list
    switches between the student's and the teacher's cutoff at every change of driver."""
    from pantea_amd.bench_models import make_nnp, student_acsf
    from pantea_amd.simulation import MDSimulator, System
    from pantea_amd.surrogate import SurrogateMD
    from pantea_amd.units import units

    s, teacher = _cfg0()
    student = make_nnp(lambda e: student_acsf(e, r_cutoff=student_rc), (16, 16), seed=99, precision="mixed")
    student.fit_scaler([s])
    sur = SurrogateMD(teacher, student, s, seed=8, judge_every=4, threshold=0.0, fallback_steps=10,
                      fallback_precision="mixed", skin=1.0)
    st = sur.run(25)
    # per window: 4 student steps rejected, 4 mixed steps rejected, 10 FP64 steps kept
    assert st.steps == 25 and st.fallback_steps == st.reference_steps == 25
    assert st.judges == st.failures == 6 and st.fallback_failures == 3
    assert st.discarded_steps == 24
    plain = System.from_structure(s, teacher, temperature=300.0, seed=8, force_mode="full", precision="float64",
                                  skin=1.0)
    MDSimulator(0.5 * units.FROM_FEMTO_SECOND, integrator="physical").run(plain, 25)
    assert float((sur.positions - plain.positions).abs().max()) < 1e-9


@pytest.mark.parametrize("student_rc", [12.0, 8.0])
def test_sampled_judge_forces_equal_full_evaluation(student_rc):
    """The sampled judge evaluates the teacher on the sampled atoms' neighbourhoods only; its
    forces on those atoms equal the full evaluation's (same kernels, same FP64 sums).  With an
    8 Bohr student the MD list has the student's cutoff and both evaluations run on the
    teacher's own list of the current positions: they equal the potential's own evaluation."""
    import torch
    from pantea_amd import water
    from pantea_amd.atoms import Structure
    from pantea_amd.bench_models import make_nnp, student_acsf, water_nnp
    from pantea_amd.surrogate import SurrogateMD

    s = Structure.from_dict(water.water_box(4096, 93.85, seed=7))
    teacher = water_nnp(s, hidden=(25, 25), seed=7, precision="float64")
    student = make_nnp(lambda e: student_acsf(e, r_cutoff=student_rc), (16, 16), seed=8, precision="mixed")
    student.fit_scaler([s])
    sur = SurrogateMD(teacher, student, s, seed=1, judge_every=10, judge_atoms=32, skin=1.0)
    md = sur.md
    assert md.nl.r_cutoff == student_rc + 1.0
    sel = np.sort(np.random.default_rng(4).choice(md.n, 32, replace=False)).astype(np.int32)
    f_sub = sur.judge_forces(sel).cpu().numpy()
    centres = sur.judge_centres
    full = torch.zeros_like(md.force)
    md.compute(full, "full", field=sur.f_judge)
    ref = full.cpu().numpy()[sel]
    print(f"sampled judge: {centres} of {md.n} centres evaluated")
    assert centres < 0.6 * md.n
    np.testing.assert_allclose(f_sub, ref, rtol=0, atol=1e-12 * np.max(np.abs(ref)))
    # the teacher's own evaluation of the same positions (input order)
    s_now = Structure.from_dict({"positions": md.positions().cpu().numpy(), "elements": s.get_elements(),
                                 "lattice": np.diag(s.box_lengths)})
    _, f_pot = teacher.energy_and_forces(s_now, mode="full", precision="float64")
    f_md = md.to_input(full).cpu().numpy()
    np.testing.assert_allclose(f_md, f_pot.cpu().numpy(), rtol=0, atol=1e-10 * np.max(np.abs(f_md)))
    st = sur.run(30)   # an untrained student: judged once at step 10, rejected, teacher to the end
    assert st.judges == st.failures == 1 and st.fallback_steps == 30
    assert all(math.isfinite(r) for r in st.judge_rmse)


@pytest.mark.parametrize("precision,tol", [("float64", 1e-12), ("mixed", 1e-5)])
def test_transposed_edge_forces_equal_reverse_index_gather(precision, tol):
    """The MD force field writes edge gradients at the reverse edges' slots
    (pt_acsf_adjoint_transposed) and gathers rows contiguously; the potential's
    evaluation uses the reverse-index gather.  Same -grad E (the FP32 edges round
    the self term differently: it is summed in FP64 before rounding)."""
    from pantea_amd.simulation import System

    s, pot = _cfg0()
    sysm = System.from_structure(s, pot, temperature=300.0, seed=1, force_mode="full", precision=precision,
                                 skin=1.0)
    f_md = sysm.forces.cpu().numpy()
    _, f_pot = pot.energy_and_forces(s, mode="full", precision=precision)
    f_pot = f_pot.cpu().numpy()
    assert np.max(np.abs(f_md - f_pot)) <= tol * np.max(np.abs(f_pot))


def test_prejudge_calls_the_judge_when_the_student_extrapolates():
    """The extrapolation count of the student's descriptors (pt_scaler_extrapolation) calls
    the judge early; with a margin no descriptor reaches, the judge keeps its interval."""
    from pantea_amd.bench_models import make_nnp, student_acsf
    from pantea_amd.simulation import BrendsenThermostat
    from pantea_amd.surrogate import SurrogateMD
    from pantea_amd.units import units

    s, teacher = _cfg0()
    student = make_nnp(student_acsf, (16, 16), seed=99, precision="mixed")
    student.fit_scaler([s])   # the start structure only: MD soon leaves its range
    runs = {}
    for margin in (0.0, 1e6):
        sur = SurrogateMD(teacher, student, s, seed=5, judge_every=20, threshold=math.inf, fallback_steps=20,
                          skin=1.0, thermostat=BrendsenThermostat(300.0, 0.5 * units.FROM_FEMTO_SECOND),
                          prejudge_margin=margin, prejudge_every=5)
        runs[margin] = sur.run(40, judge_tail=True).as_dict()
        print(margin, runs[margin])
    assert runs[0.0]["prejudge_triggers"] > 0 and runs[0.0]["extrapolated_rows"] > 0
    assert runs[0.0]["judges"] > runs[1e6]["judges"]
    assert runs[1e6]["prejudge_triggers"] == 0 and runs[1e6]["judges"] == 2
    for r in runs.values():   # the threshold passes every judge: the student drives all
        assert r["failures"] == 0 and r["student_kept_steps"] == 40 and r["steps"] == 40
