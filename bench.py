#!/usr/bin/env python
"""Benchmark: atom-steps/s of 98,304-atom water NNP-MD on one MI355X.

Workload: 32,768 H2O (98,304 atoms) in a 187.7 Bohr periodic box, velocity
Verlet with a = F/m, F = -grad E, dt = 0.5 fs, Berendsen 300 K, skin
neighbour list.  Teacher NNP: 52 ACSF per element (tanh cutoff 12 Bohr),
MLP 52-64-64-1 tanh, weights ~ N(0, 1/fan_in), min-max scaler fitted on the
structure.  Every step runs without a host read (skin check and rebuild on
the device) and is replayed from a HIP graph.  Legs, all on the same
structure and protocol:

  surrogate (headline): the trained student (34 ACSF, MLP 34-16-16-1, mixed
      precision) drives; the FP64 teacher judges every 20 steps on the
      neighbourhoods of 128 sampled atoms (every 5 steps while the student is
      on probation: until its first pass and after a failure); a force RMSE
      above 5 meV/Bohr rewinds to the last good state and the
      mixed-precision teacher drives the next 200 steps, judged the same way
      (a failure there hands the window to the FP64 teacher).
  surrogate at the student's accuracy: the same with the threshold at twice
      the student's force RMSE, i.e. what the learned part buys when it drives.
  teacher mixed: descriptors + adjoint FP32, MLP split-BF16 on MFMA, FP64
      energy sum, force gather and integrator.
  teacher FP64: the plain FP64 GPU path.
  student alone: the student driving without a judge.
CPU baselines: the FP64 oracle with a cell list (headline structure and two
smaller boxes, sampled centres) and the reference's dense algorithm on an
N-ladder.

One step = position update, neighbour maintenance, one full energy + force
evaluation, velocity update and thermostat; inputs are resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path
from types import SimpleNamespace

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# MI355X published peaks
PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_FP64_MFMA_TFLOPS = 78.6   # MI355X dense FP64 matrix
kWaveThreads = 64   # a descriptor launch runs one wavefront per centre


def _log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def _sync():
    import torch

    torch.cuda.synchronize()


def neighbour_stats(s, rc, sample=256, seed=0):
    """Mean neighbours and live triplets (pairs {j,k} with r_jk < rc) per atom, from a sample of centres."""
    pos = s.positions.cpu().numpy()
    L = np.asarray(s.box_lengths)
    rng = np.random.default_rng(seed)
    idx = rng.choice(s.natoms, min(sample, s.natoms), replace=False)
    n_nb, n_tr = [], []
    for i in idx:
        d = pos[i] - pos
        d -= L * np.round(d / L)
        r = np.linalg.norm(d, axis=1)
        m = (r <= rc) & (r > 0)
        R = d[m]
        n_nb.append(R.shape[0])
        djk = R[:, None, :] - R[None, :, :]
        djk -= L * np.round(djk / L)
        rjk = np.linalg.norm(djk, axis=2)
        iu = np.triu_indices(R.shape[0], 1)
        n_tr.append(int(np.count_nonzero(rjk[iu] < rc)))
    return float(np.mean(n_nb)), float(np.mean(n_tr))


def algorithmic_ops(n_nb, n_trip, g2=8, g3=12, n_eta=3, quartic=True, angular_groups_per_neighbour=2):
    """Algorithmic ops per centre of one descriptor forward and adjoint pass
    (add/mul/compare 1, FMA 2, transcendental 1; screening, queueing, index and
    LDS-address arithmetic are not counted).

    forward (the per-pair math of the descriptor definitions):
        (15 + 6 g2) n_nb + (26 + 5 g3) n_trip
    adjoint (counted from the pair and neighbour bodies of k_acsf in adjoint mode):
      per live pair {j, k} (r_jk < rc), all functions of its angular group:
        geometry 31   (r_ij^2, r_ik^2 10; R_jk 3; r_jk^2 5; 1/sqrt + r 2; cos 7;
                       S 2; 1/r_ij^2, 1/r_ik^2 2)
        cutoff 18     (f and f' of r_jk 10; F = fc fc fc and its three partials 8)
        angular       quartic class (zeta {1,4} x lambda +-1): shared powers of cos 8,
                      then 18 per eta (exp 2, U 6, U' 4, P 1, the three sums 5);
                      two-term zeta = 1 class: terms 2, then 16 per eta
        assembly 17   (d/dr_jk, d/dcos, the two radial coefficients, 1/(r_ij r_ik) term)
        gradient 36   (two 3-vectors of three terms 30, six accumulations 6)
        -> 110 + 18 n_eta (quartic) or 104 + 16 n_eta (two-term)
      per neighbour: radial group 18 + 11 g2 (r 1, f and f' 10, per function d,
        exp, derivative and weight 11, scaling and three accumulations 7), the
        angular groups' f and f'/r 12 each, the centre sum 3
        -> 21 + 11 g2 + 12 * angular_groups_per_neighbour
    """
    fwd = (15.0 + 6.0 * g2) * n_nb + (26.0 + 5.0 * g3) * n_trip
    pair = (110.0 + 18.0 * n_eta) if quartic else (104.0 + 16.0 * n_eta)
    nb = 21.0 + 11.0 * g2 + 12.0 * angular_groups_per_neighbour
    adj = nb * n_nb + pair * n_trip
    return fwd, adj


def profiled_traffic(kernel: str, grid: int):
    """HBM bytes per launch of ``kernel`` at ``grid`` work-items (mean over those dispatches) from the
    committed PMC summary."""
    path = ROOT / "profiles" / "hbm_traffic_water98k_md.md"
    if not path.exists():
        return None
    vals = []
    for line in path.read_text().splitlines():
        if line.startswith(f"| `{kernel}` |"):
            cols = [c.strip() for c in line.strip("|").split("|")]
            if int(cols[1]) == grid:   # the dispatches of this launch size
                vals.append((float(cols[5]) + float(cols[6])) * 1e6)
    return float(np.mean(vals)) if vals else None


def _oracle_threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def _mlp_weights(onn, types, centres, G):
    """w = dE_i/dG_i of the oracle's scaler + MLP for the rows G of ``centres``."""
    from oracle import pyoracle

    w = np.zeros_like(G)
    for e in onn.elements:
        sel = types[centres] == onn.type_of[e]
        if not np.any(sel):
            continue
        x = pyoracle.scaler_transform(onn.scaler_kind, onn.scaler_params[e], G[sel], onn.smin, onn.smax)
        _, dy = pyoracle.mlp_forward_backward(onn.layers[e], x)
        w[sel, : dy.shape[1]] = dy * pyoracle.scaler_slope(onn.scaler_kind, onn.scaler_params[e], onn.smin,
                                                           onn.smax)
    return w


def _host_water(pot, n_molecules: int, box: float, seed: int = 2023):
    """(oracle NNP, wrapped positions, types, box) of a water box, on the host only."""
    from types import SimpleNamespace

    from oracle import pyoracle
    from oracle.bridge import oracle_nnp
    from pantea_amd import water
    from pantea_amd.atoms import ElementMap

    d = water.water_box(n_molecules, box, seed=seed)
    emap = ElementMap.from_list(d["elements"])
    onn = oracle_nnp(pot, SimpleNamespace(element_map=emap))
    L = np.array([box, box, box], dtype=np.float64)
    types = np.array([emap.element_to_atom_type[e] for e in d["elements"]], dtype=np.int32)
    return onn, pyoracle.wrap_positions(d["positions"], L), types, L


def cpu_celllist(onn, pos, types, L, rc: float, n_centres: int):
    """Cell-list CPU port (FP64 oracle, OpenMP): cell-list neighbour rows of all
    atoms, then for a fixed sample of centres the descriptors (per-function loops
    over the centre's neighbour pairs), scaler + MLP and the force VJP.
    atom-steps/s = 1 / (list time per atom + centre time per centre)."""
    from oracle import pyoracle

    n = len(types)
    t0 = time.perf_counter()
    row_off, idx = pyoracle.cell_neighbors(pos, L, rc)
    t_list = time.perf_counter() - t0
    k = min(n_centres, n)
    centres = np.sort(np.random.default_rng(1).choice(n, k, replace=False)).astype(np.int32)
    t0 = time.perf_counter()
    for e in onn.elements:
        sel = centres[types[centres] == onn.type_of[e]]
        if sel.size == 0:
            continue
        G, _ = pyoracle.celllist_centres(pos, types, L, row_off, idx, onn.tables[e], sel)
        w = _mlp_weights(onn, types, sel, G)
        pyoracle.celllist_centres(pos, types, L, row_off, idx, onn.tables[e], sel, w)
    t_c = time.perf_counter() - t0
    value = 1.0 / (t_list / n + t_c / k)
    return value, {"atoms": n, "centres": k, "list_seconds": t_list, "centre_seconds": t_c}


def cpu_reference_ladder(pot, sizes=(32, 64, 128), seed=2023):
    """The reference's dense algorithm (all N atoms per radial function, all N x N
    ordered pairs per angular function, every term evaluated and masked) on water
    boxes of 3n atoms at the 128-H2O density: descriptors of every atom, scaler +
    MLP, then the self-term force pass through the same loops (jax.grad w.r.t. the
    central positions, force.py:16-43).  Returns per-N atom-steps/s and the fitted
    exponent of time in N."""
    from oracle import pyoracle

    rho = 128 / 29.56 ** 3
    out = []
    for nm in sizes:
        L = (nm / rho) ** (1.0 / 3.0)
        onn, pos, types, box = _host_water(pot, nm, L, seed)
        st = SimpleNamespace(natoms=len(types), box_lengths=box)
        t0 = time.perf_counter()
        for e in onn.elements:
            sel = np.nonzero(types == onn.type_of[e])[0].astype(np.int32)
            G, _ = pyoracle.reference_algorithm(pos, types, st.box_lengths, onn.tables[e], sel)
            w = _mlp_weights(onn, types, sel, G)
            pyoracle.reference_algorithm(pos, types, st.box_lengths, onn.tables[e], sel, w)
        t = time.perf_counter() - t0
        out.append({"atoms": st.natoms, "box_bohr": L, "seconds_per_step": t, "atom_steps_per_s": st.natoms / t})
    n = np.array([o["atoms"] for o in out], dtype=np.float64)
    t = np.array([o["seconds_per_step"] for o in out])
    p = float(np.polyfit(np.log(n), np.log(t), 1)[0])
    return out, p


def cpu_baselines(pot, molecules: int, box: float, threads: int):
    """All CPU legs: the headline object (the cell-list port on the headline
    structure) plus the cell-list port at 128 and 4096 H2O and the
    reference-algorithm ladder."""
    legs = {}
    head = None
    for nm, L in ((128, 29.56), (4096, 93.85), (molecules, box)):
        onn, pos, types, Lv = _host_water(pot, nm, L)
        v, inf = cpu_celllist(onn, pos, types, Lv, pot.r_cutoff, 2048)
        legs[f"{nm} H2O"] = {"atom_steps_per_s": v, **inf}
        head = (v, inf)
    value, info = head
    cpu = {"value": value, "unit": "atom-steps/s", "cores": threads, "kind": "port",
           "sample": f"FP64 oracle (C, OpenMP) with a cell list on the {info['atoms']}-atom structure: neighbour rows "
                     f"of all atoms ({info['list_seconds']:.2f} s), then descriptors (per-function loops over each "
                     f"centre's neighbour pairs), scaler + MLP and the force VJP for a fixed sample of "
                     f"{info['centres']} centres ({info['centre_seconds']:.2f} s); atom-steps/s = 1 / (list time per "
                     f"atom + centre time per centre)"}
    ladder, expo = cpu_reference_ladder(pot)
    last = ladder[-1]
    n_head = info["atoms"]
    extrap = last["atom_steps_per_s"] * (last["atoms"] / n_head) ** (expo - 1.0)
    ref = {"kind": "reference algorithm (dense: all N atoms per radial function, all N x N pairs per angular "
                   "function, every term evaluated and masked)", "cores": threads, "ladder": ladder,
           "time_exponent_fit": expo,
           "extrapolated_atom_steps_per_s_at_headline_atoms": extrap,
           "note": f"measured at {', '.join(str(o['atoms']) for o in ladder)} atoms; the value at {n_head} atoms "
                   "is an extrapolation of the fitted power law (time ~ N^p), not a measurement"}
    return cpu, {"celllist_port": legs, "reference_algorithm": ref}


def _thermostat(args):
    from pantea_amd.simulation import BrendsenThermostat
    from pantea_amd.units import units

    return BrendsenThermostat(300.0, args.tau_fs * units.FROM_FEMTO_SECOND) if args.tau_fs > 0 else None


def run_md(pot, s, precision, steps, warmup, args, instrument=False):
    """Plain MD of one potential; returns (seconds, per-phase HIP-event timers or None, system).

    Timed runs are not instrumented (the phase events cost host time and
    GPU idle); an instrumented run gives the per-phase breakdown."""
    from pantea_amd.simulation import MDSimulator, System
    from pantea_amd.units import units

    sysm = System.from_structure(s, pot, temperature=300.0, seed=2023, force_mode="full",
                                 precision=precision, skin=args.skin, graph=not args.no_graph and not instrument)
    md = MDSimulator(time_step=0.5 * units.FROM_FEMTO_SECOND, thermostat=_thermostat(args), integrator="physical")
    md.run(sysm, warmup)
    _sync()
    eng = sysm._engine()
    eng.timers = {} if instrument else None
    t0 = time.perf_counter()
    md.run(sysm, steps)
    _sync()
    dt = time.perf_counter() - t0
    timers, eng.timers = eng.timers, None
    return dt, timers, sysm


def run_surrogate(teacher, student, s, steps, warmup, args, instrument=False, threshold_mev=None):
    from pantea_amd.surrogate import SurrogateMD
    from pantea_amd.units import units

    thr = args.threshold_mev if threshold_mev is None else threshold_mev
    sur = SurrogateMD(teacher, student, s, seed=2023, judge_every=args.judge_every,
                      threshold=thr / units.TO_MILLI_ELECTRON_VOLT,
                      fallback_steps=args.fallback_steps, fallback_precision=args.fallback_precision,
                      skin=args.skin, thermostat=_thermostat(args), judge_atoms=args.judge_atoms or None,
                      probation_every=args.probation_every, graph=not args.no_graph and not instrument)
    sur.run(warmup)
    _sync()
    before = sur.stats.as_dict()
    n_rmse = len(sur.stats.judge_rmse)
    sur.md.timers = {} if instrument else None
    t0 = time.perf_counter()
    sur.run(steps)
    _sync()
    dt = time.perf_counter() - t0
    timers, sur.md.timers = sur.md.timers, None
    after = sur.stats.as_dict()
    delta = {k: after[k] - before[k] for k in ("steps", "student_steps", "fallback_steps", "reference_steps",
                                                "discarded_steps", "judges", "failures", "fallback_failures",
                                                "student_kept_steps")}
    r = np.asarray(sur.stats.judge_rmse[n_rmse:]) * units.TO_MILLI_ELECTRON_VOLT
    who = np.asarray(sur.stats.judge_driver[n_rmse:])
    delta["fallback_precision"] = args.fallback_precision
    # share of the kept trajectory driven by the teacher (a rewind can reach back into the warm-up)
    delta["teacher_share"] = after["teacher_share"]
    rs, rf = r[who == "student"], r[who == "fallback"]
    delta["student_judges"] = int(rs.size)
    delta["student_judge_rmse_mev_per_bohr_median"] = float(np.median(rs)) if rs.size else None
    delta["student_judge_rmse_mev_per_bohr_min"] = float(np.min(rs)) if rs.size else None
    delta["fallback_judges"] = int(rf.size)
    delta["fallback_judge_rmse_mev_per_bohr_max"] = float(np.max(rf)) if rf.size else None
    delta["threshold_mev_per_bohr"] = thr
    delta["judge_centres"] = getattr(sur, "judge_centres", None)
    return dt, timers, delta


def kalman_flops(n_w: int, m: int) -> float:
    """Algorithmic FP64 flops of one pt_kalman_update in the form it computes.  m <= n_w: X = P H
    (2 n_w^2 m), A = H^T X (2 m^2 n_w), Cholesky of A (m^3 / 3), the solve for Y = A^-1 X^T (2 m^2 n_w),
    P -= Y^T X^T (2 n_w^2 m).  m > n_w (n_w x n_w factors): Cholesky of P (n_w^3 / 3), G = L^T H
    (n_w^2 m), S = G G^T (n_w^2 m), Cholesky of S (n_w^3 / 3), M = L Ls^-T (n_w^3), M M^T (n_w^3)."""
    if m > n_w:
        return 2.0 * n_w * n_w * m + 8.0 / 3.0 * n_w ** 3
    return 4.0 * n_w * n_w * m + 4.0 * m * m * n_w + m ** 3 / 3.0


def run_kalman(args):
    """The Kalman-filter trainer's per-structure cost on the student network (34-16-16-1 per element):
    one force update of a 512-molecule box (1,536 atoms, m = 4,608 observations) and one energy update,
    device time per call from HIP events; plus the same force update on the host with numpy (the
    reference's formula, kalman_filter.py:196-233) as the CPU figure."""
    import numpy as np
    import torch

    from pantea_amd import bench_models, water
    from pantea_amd.atoms import Structure
    from pantea_amd.kalman import KalmanFilter
    from pantea_amd.potential import NNPSettings

    s = Structure.from_dict(water.water_box(512, 187.7 / 4, seed=5))
    student = bench_models.water_student(None, seed=12, precision="float64")
    student.fit_scaler([s])
    E, F = student.energy_and_forces(s, mode="reference", precision="float64")
    s.forces = F + 0.01 * torch.sin(s.positions)
    s.total_energy = E + 0.5
    settings = NNPSettings(elements=list(student.elements), global_nodes_short=[16, 16],
                           global_activation_short=["tanh", "tanh", "identity"], symfunction_short=[])
    kf = KalmanFilter.from_settings(student, settings)
    kf.dev = torch.device("cuda")
    kf._bind()
    items = kf._prepare([s])
    kf._buffers(items)
    it = items[0]
    n_w = kf.num_states

    def timed(fn, reps):
        fn()
        _sync()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        _sync()
        return a.elapsed_time(b) / reps

    reps = max(1, args.kalman_updates)
    jac_ms = timed(lambda: kf.calculate_loss_force(it), reps)
    xi, H = kf.calculate_loss_force(it)
    xi, H = xi.clone(), H.clone()
    m = int(xi.numel())
    upd_ms = timed(lambda: kf.update_state_vector(xi, H), reps)
    xe, He = kf.calculate_loss_energy(it)
    xe, He = xe.clone(), He.clone()
    en_ms = timed(lambda: (kf.calculate_loss_energy(it), kf.update_state_vector(xe, He)), reps)
    if int(kf.info_sum.item()) != 0:
        raise RuntimeError("Kalman update: a Cholesky factorisation failed")
    flops = kalman_flops(n_w, m)
    out = {"network": "student 34-16-16-1 per element (H, O)", "atoms": s.natoms, "n_weights": n_w,
           "observations": m, "force_jacobian_ms": jac_ms, "force_update_ms": upd_ms,
           "energy_observation_and_update_ms": en_ms, "update_flops": flops,
           "update_tflops": flops / (upd_ms * 1e-3) / 1e12,
           "update_frac_of_fp64_peak": flops / (upd_ms * 1e-3) / 1e12 / PEAK_FP64_MFMA_TFLOPS,
           "update_form": "n_w x n_w factors (m > n_w)" if m > n_w else "m x m factor",
           "note": "flops of the form computed (kalman_flops in bench.py: m > n_w: 2 n_w^2 m + 8/3 n_w^3, else "
                   "4 n_w^2 m + 4 m^2 n_w + m^3/3); its two Cholesky factorisations of n_w rows (blocked, one "
                   "wavefront per 64 x 64 diagonal block) and their many small BLAS calls are latency-bound; "
                   f"peak = MI355X FP64 matrix {PEAK_FP64_MFMA_TFLOPS} TFLOP/s; HIP events over {reps} calls"}
    if not args.no_cpu_baseline:
        # the same update in numpy on the host, the reference's explicit inverse (one call)
        P0 = kf.P.cpu().numpy().T.copy()
        Hn = H.cpu().numpy().T.copy()         # n_w x m
        xn = xi.cpu().numpy()
        t0 = time.perf_counter()
        X = P0 @ Hn
        A = Hn.T @ X + (1.0 / kf.eta) * np.eye(m)
        K = X @ np.linalg.inv(A)
        P1 = P0 - K @ X.T + kf.q * np.eye(n_w)
        W1 = K @ xn
        out["cpu_numpy_force_update_ms"] = (time.perf_counter() - t0) * 1e3
        out["cpu_threads"] = _oracle_threads()
        del P1, W1
    return out


def phase_table(timers):
    """{phase: (total ms, launch groups)} from HIP-event pairs."""
    out = {}
    for ph, evs in timers.items():
        ts = [a.elapsed_time(b) for a, b in evs if b is not None]
        out[ph] = (float(np.sum(ts)), len(ts))
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400, help="timed steps of the headline (surrogate) leg")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--teacher-steps", type=int, default=100, help="timed steps of the mixed teacher leg")
    ap.add_argument("--fp64-steps", type=int, default=60)
    ap.add_argument("--sweep-steps", type=int, default=1000,
                    help="timed steps of the surrogate at the student's own accuracy level (0: skip)")
    ap.add_argument("--skin", type=float, default=0.0,
                    help="neighbour-list skin in Bohr (0: rebuild every step, measured fastest for this teacher)")
    ap.add_argument("--tau-fs", type=float, default=0.5, help="Berendsen time constant (0: NVE)")
    ap.add_argument("--judge-every", type=int, default=20)
    ap.add_argument("--fallback-steps", type=int, default=200)
    ap.add_argument("--probation-every", type=int, default=5,
                    help="judge interval of a student on probation (not yet passed, or failed), until it passes")
    ap.add_argument("--fallback-precision", default="mixed", choices=("mixed", "float64"),
                    help="precision of the fallback teacher (below the judge's it is judged too)")
    ap.add_argument("--threshold-mev", type=float, default=5.0)
    ap.add_argument("--judge-atoms", type=int, default=128,
                    help="judge on this many sampled atoms (their neighbourhoods only); 0: every atom")
    ap.add_argument("--molecules", type=int, default=32768)
    ap.add_argument("--box", type=float, default=187.7)
    ap.add_argument("--student", default=None, help="student directory (default: the packaged water student)")
    ap.add_argument("--no-graph", action="store_true", help="launch every step from the host (no HIP graphs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--kalman-updates", type=int, default=3,
                    help="timed calls of the Kalman-filter leg (force Jacobian, force update, energy update); 0: skip")
    ap.add_argument("--kalman-only", action="store_true",
                    help="run only the Kalman-filter leg and print its JSON (for profiling)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="after the run, launch 3 plain FP64 copy kernels of MIB MiB each (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus N > 1 is not supported)", file=sys.stderr)
        return 2

    import torch

    from pantea_amd import bench_models
    from pantea_amd.student import ASSET_DIR
    from pantea_amd.units import units

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.kalman_only:
        print(json.dumps({"kalman": run_kalman(args)}), flush=True)
        return 0
    mev = units.TO_MILLI_ELECTRON_VOLT
    t_setup = time.perf_counter()
    s, teacher = bench_models.water_teacher(precision="mixed", molecules=args.molecules, box=args.box)
    n = s.natoms
    _sync()
    _log(f"[bench] {n} atoms, box {args.box} Bohr, teacher ready ({time.perf_counter() - t_setup:.1f} s)")
    student_dir = Path(args.student) if args.student else ASSET_DIR
    student = bench_models.water_student(student_dir) if (student_dir / "weights.001.npz").exists() else None
    if student is None:
        _log(f"[bench] no trained student under {student_dir}: the headline falls back to the mixed teacher")

    # accuracy of the mixed teacher vs the FP64 GPU path (itself checked against the oracle in tests)
    e64, f64 = teacher.energy_and_forces(s, mode="full", precision="float64")
    em, fm = teacher.energy_and_forces(s, mode="full", precision="mixed")
    _sync()
    e_err = abs(float(em) - float(e64)) / n * mev
    f_rmse = float(torch.sqrt(torch.mean((fm - f64) ** 2))) * mev
    f_rms = float(torch.sqrt(torch.mean(f64 ** 2))) * mev
    _log(f"[bench] mixed teacher vs FP64: energy {e_err:.4g} meV/atom, force RMSE {f_rmse:.4g} meV/Bohr "
         f"(|F| rms {f_rms:.4g})")
    stud_acc = None
    if student is not None:
        es, fs = student.energy_and_forces(s, mode="full", precision="mixed")
        _sync()
        stud_acc = {"energy_error_mev_per_atom": abs(float(es) - float(e64)) / n * mev,
                    "force_rmse_mev_per_bohr": float(torch.sqrt(torch.mean((fs - f64) ** 2))) * mev,
                    "r_cutoff_bohr": student.r_cutoff}
        metrics = student_dir / "metrics.json"
        if metrics.exists():
            val = json.loads(metrics.read_text()).get("validation_metrics", {})
            stud_acc["validation_force_rmse_mev_per_bohr"] = val.get("force_rmse_mev_per_bohr")
        _log(f"[bench] student vs FP64 teacher on the start structure: {stud_acc}")

    # teacher legs
    dt_m, _, _ = run_md(teacher, s, "mixed", args.teacher_steps, args.warmup, args)
    mixed_value = n * args.teacher_steps / dt_m
    _log(f"[bench] teacher mixed MD: {1e3 * dt_m / args.teacher_steps:.3f} ms/step, {mixed_value:.4g} atom-steps/s")
    phase_steps_m = min(args.teacher_steps, 50)
    _, timers_m, _ = run_md(teacher, s, "mixed", phase_steps_m, 2, args, instrument=True)
    phases_m = phase_table(timers_m)
    fp64_value = None
    if not args.no_fp64:
        dt64, _, _ = run_md(teacher, s, "float64", args.fp64_steps, 1, args)
        fp64_value = n * args.fp64_steps / dt64
        _log(f"[bench] teacher FP64 MD: {1e3 * dt64 / args.fp64_steps:.3f} ms/step, {fp64_value:.4g} atom-steps/s")

    # the student alone (no judge): the rate the surrogate reaches when every judge passes
    student_value = None
    if student is not None:
        dt_st, _, _ = run_md(student, s, "mixed", args.teacher_steps, args.warmup, args)
        student_value = n * args.teacher_steps / dt_st
        _log(f"[bench] student mixed MD (no judge): {1e3 * dt_st / args.teacher_steps:.3f} ms/step, "
             f"{student_value:.4g} atom-steps/s")

    # headline leg: the configured surrogate (threshold as given)
    sur = None
    sweep = None
    if student is not None:
        dt_s, _, sur = run_surrogate(teacher, student, s, args.steps, args.warmup, args)
        value = n * args.steps / dt_s
        ms_per_step = 1e3 * dt_s / args.steps
        steps = args.steps
        _log(f"[bench] surrogate MD: {ms_per_step:.3f} ms/step, {value:.4g} atom-steps/s, {sur}")
        phase_steps = min(args.steps, 200)
        _, timers_s, _ = run_surrogate(teacher, student, s, phase_steps, args.warmup, args, instrument=True)
        phases = phase_table(timers_s)
        # the same surrogate with the judge's threshold at the student's own accuracy level: twice
        # its force RMSE on held-out teacher MD snapshots (the training's validation set; else on
        # the start structure): what the learned part buys when it drives
        if args.sweep_steps > 0 and stud_acc is not None:
            basis = "validation" if stud_acc.get("validation_force_rmse_mev_per_bohr") else "start structure"
            thr = 2.0 * (stud_acc.get("validation_force_rmse_mev_per_bohr") or stud_acc["force_rmse_mev_per_bohr"])
            dt_w, _, sw = run_surrogate(teacher, student, s, args.sweep_steps, args.warmup, args, threshold_mev=thr)
            sweep = {"threshold_mev_per_bohr": thr, "threshold_basis": f"2 x the student's force RMSE ({basis})",
                     "steps": args.sweep_steps,
                     "atom_steps_per_s": n * args.sweep_steps / dt_w, "ms_per_step": 1e3 * dt_w / args.sweep_steps,
                     "over_fp64": (n * args.sweep_steps / dt_w / fp64_value) if fp64_value else None,
                     "surrogate": sw}
            _log(f"[bench] surrogate at {thr:.4g} meV/Bohr: {sweep['ms_per_step']:.3f} ms/step, {sw}")
    else:
        value, ms_per_step, phases, steps = mixed_value, 1e3 * dt_m / args.teacher_steps, phases_m, args.teacher_steps
        phase_steps = phase_steps_m

    # roofline of the descriptor kernel with the most time in the headline run
    n_nb, n_trip = neighbour_stats(s, teacher.r_cutoff)
    cands = {k: v for k, v in phases.items() if k.rsplit("/", 1)[-1] in ("acsf_forward", "acsf_adjoint")}
    dom = max(cands, key=lambda k: cands[k][0])
    label, ph = (dom.rsplit("/", 1) if "/" in dom else ("teacher-mixed", dom))
    is_student = label.startswith("student")
    fp64 = label.endswith("float64")
    pot_of = student if is_student else teacher
    if is_student:
        nb_s, tr_s = neighbour_stats(s, student.r_cutoff)
        # the student's tables are completed to the four-term class: quartic adjoint
        fwd_ops, adj_ops = algorithmic_ops(nb_s, tr_s, 8, 6, quartic=True)
    else:
        fwd_ops, adj_ops = algorithmic_ops(n_nb, n_trip, 8, 12, quartic=True)
    mode = 1 if ph == "acsf_adjoint" else 0
    total_ms, groups = phases[dom]
    # elements with byte-identical descriptor tables share one launch per evaluation
    n_launch = len({bytes(pot_of.atomic_potentials[e].descriptor.device_table(s.element_map, s.device).host)
                    for e in pot_of.elements})
    launch_ms = total_ms / (n_launch * groups)
    launch_ops = (adj_ops if mode else fwd_ops) * n / n_launch
    achieved = launch_ops / (launch_ms * 1e-3) / 1e12 if launch_ms > 0 else 0.0
    peak = PEAK_FP64_TFLOPS if fp64 else PEAK_FP32_TFLOPS
    kind = pot_of.atomic_potentials["O"].descriptor.device_table(s.element_map, s.device).kind
    if mode == 0 and fp64 and kind != 0:
        variant = "_w4"
    elif mode == 0 and not fp64 and kind in (2, 3):
        variant = "_w8"
    else:
        variant = ""
    kname = f"k_acsf{variant}<{'double' if fp64 else 'float'}, {mode}, {kind}>"
    roofline = {
        "bound": "valu",
        "achieved": achieved,
        "peak": peak,
        "unit": "TFLOP/s",
        "frac": achieved / peak,
        "traffic": profiled_traffic(kname, int(kWaveThreads * n / n_launch)),
        "kernel": f"{kname} ({'descriptor adjoint' if mode else 'descriptor forward'}, {label})",
        "avg_launch_ms": launch_ms,
        "algorithmic_ops_per_launch": launch_ops,
        "note": "vector-ALU bound (no matrix cores in this kernel): algorithmic ops per centre, forward "
                "(15 + 6 g2) n_nb + (26 + 5 g3) n_trip, adjoint counted from its neighbour and pair bodies "
                "(algorithmic_ops in bench.py); transcendental = 1, FMA = 2; "
                f"measured n_nb={n_nb:.1f}, n_trip={n_trip:.0f} per atom at the teacher's cutoff; a launch covers "
                f"{'all' if n_launch == 1 else f'1/{n_launch} of the'} centres; average launch time from HIP events "
                f"over {groups * n_launch} launches of an instrumented {phase_steps}-step repeat of the headline run; "
                f"peak = published MI355X {'FP64' if fp64 else 'FP32'} vector; traffic = HBM bytes per launch from "
                "the FETCH_SIZE/WRITE_SIZE profile, copy-kernel calibrated",
    }
    _log(f"[bench] roofline: {roofline['kernel']} {achieved:.3g} TFLOP/s = {roofline['frac']:.3f} of peak")

    cpu, cpu_legs = None, None
    if not args.no_cpu_baseline:
        try:
            cpu, cpu_legs = cpu_baselines(teacher, args.molecules, args.box, _oracle_threads())
            _log(f"[bench] cpu baseline: {cpu['value']:.4g} atom-steps/s on {cpu['cores']} threads")
        except Exception as exc:  # the baseline is a report, not the measurement
            _log(f"[bench] cpu baseline failed: {exc!r}")

    kalman = None
    if args.kalman_updates > 0:
        kalman = run_kalman(args)
        _log(f"[bench] Kalman filter: {kalman}")

    if args.copy_probe > 0:
        # y = 1.0 * x streams exactly 2 * MIB MiB through HBM per launch
        src = torch.rand(args.copy_probe * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        _sync()
        _log(f"[bench] copy probe: 3 x {args.copy_probe} MiB read + {args.copy_probe} MiB write")
        del src, dst

    thermo = f"Berendsen 300 K tau {args.tau_fs} fs" if args.tau_fs > 0 else "NVE"
    judged = f"{args.judge_atoms} sampled atoms" if args.judge_atoms else "all atoms"
    fb = ("mixed-precision teacher steps (judged the same way; failing -> FP64 teacher)"
          if args.fallback_precision == "mixed" else "FP64 teacher steps")
    drove = bool(sur and sur["student_kept_steps"] > 0)   # steps of the student that no rewind discarded
    if sur is None:
        head = "mixed-precision teacher (no trained student found)"
    else:
        head = (f"surrogate: student 34 ACSF (rc {student.r_cutoff} Bohr) 34-16-16-1 mixed drives, FP64 teacher "
                f"judges every {args.judge_every} steps (a student on probation, until it passes: every "
                f"{args.probation_every}) on {judged}, > {args.threshold_mev} meV/Bohr -> rewind + "
                f"{args.fallback_steps} {fb}")
        if not drove:
            head += ("; in the timed window no step of the student was kept (it fails the judge), so the value is the "
                     "judged mixed-precision teacher's rate: the surrogate's own rate at this threshold is "
                     "UNMEASURED (see surrogate_at_student_accuracy)")
    out = {
        "metric": "atom-steps/s, 98k-atom water NNP-MD: mixed+surrogate vs FP64; force RMSE",
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+bf16x3(mfma)+f64-accumulate; f64 judge",
        "data": "synthetic water (seeded random sequential addition), random-init teacher weights, "
                "student trained on the GPU on teacher MD data",
        "config": {"workload": f"water NNP-MD, {n} atoms ({n // 3} H2O), box {args.box} Bohr, rc 12 Bohr, "
                               f"teacher 52 ACSF/element + MLP 52-64-64-1, dt 0.5 fs, {thermo}; headline: {head}",
                   "atoms": n, "skin_bohr": args.skin, "parallelism": "single GPU",
                   "step_launch": "host-free step (device skin check + gated rebuild)"
                                  + ("" if args.no_graph else ", replayed from a HIP graph")},
        "headline_driver": ("student" if drove else "teacher (mixed precision, judged)") if sur else "teacher",
        "fp64_atom_steps_per_s": fp64_value,
        "mixed_teacher_atom_steps_per_s": mixed_value,
        "student_mixed_atom_steps_per_s": student_value,
        "student_over_fp64": (student_value / fp64_value) if (student_value and fp64_value) else None,
        "value_over_fp64": (value / fp64_value) if fp64_value else None,
        "mixed_teacher_over_fp64": (mixed_value / fp64_value) if fp64_value else None,
        "surrogate": sur,
        "surrogate_at_student_accuracy": sweep,
        "student_accuracy_vs_fp64": stud_acc,
        "energy_error_mev_per_atom": e_err,
        "force_rmse_mev_per_bohr": f_rmse,
        "accuracy_note": "energy_error / force_rmse: mixed teacher vs FP64 teacher on the start structure; "
                         "surrogate.*_judge_rmse_*: the driver (student / mixed fallback) vs the FP64 "
                         "teacher at its judges",
        "force_rms_mev_per_bohr": f_rms,
        "threshold_over_force_rms": args.threshold_mev / f_rms,
        "phase_ms_per_step": {k: v[0] / phase_steps for k, v in phases.items()},
        "phase_note": f"HIP events of instrumented repeats ({phase_steps} headline steps, {phase_steps_m} mixed "
                      "teacher steps, launched from the host); the timed runs carry no events",
        "teacher_mixed_phase_ms_per_step": {k: v[0] / phase_steps_m for k, v in phases_m.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_baseline_legs": cpu_legs,
        "kalman": kalman,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
