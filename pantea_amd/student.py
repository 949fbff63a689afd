"""Student NNP trained on the GPU from teacher trajectories.

Pipeline:
  1. ``teacher_snapshots``: teacher MD (Berendsen, 300 K) on water boxes of
     uneven size at 1 g/cm^3; every ``every`` steps the configuration is
     labelled with FP64 teacher energies and forces (F = -grad E).
  2. ``StudentTrainer``: per snapshot the student's neighbour list and
     descriptors are built once (positions are fixed); every optimisation
     step then runs, per structure,
        MLP energy + dE/dG          pt_mlp_energy_grad
        E = sum_i E_i               pt_reduce_sum_f64
        F = -J^T dE/dG              pt_acsf_adjoint + pt_force_gather
        delta = dL/dF, dL/dE        pt_train_seeds
        T = J delta                 pt_acsf_tangent (forward mode)
        dL/dtheta accumulation      pt_mlp_train_grad (forward-over-reverse)
     and after each batch one Adam step (pt_adam_step) and a repack.
Loss per structure: w_E ((E - E*)/N)^2 + w_F / (3N) sum |F - F*|^2,
averaged over the batch.  The reference's own trainer
(potentials/nnp/gradient_descent.py:157-233) is broken and never run by its
tests; its loss (energy MSE per atom + beta^2 force MSE) and optax Adam
defaults (settings.py:116-121) are followed where they apply.
"""

from __future__ import annotations

import json
import math
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .atoms import Structure
from .engine import NeighborList, Reducer, acsf_adjoint, acsf_forward, acsf_tangent, force_gather, gather_rows
from .potential import PRECISION_MODES, NeuralNetworkPotential
from .units import units


@dataclass
class Snapshot:
    """One labelled configuration, resident on the device in internal order."""

    n: int
    nl: NeighborList
    blocks: List[Tuple[str, int, int]]   # (element, begin, count) in internal order
    G: object                            # (n, ld) student descriptors
    F_ref: object                        # (n, 3) FP64 teacher forces, internal order
    E_ref: float
    emap_key: tuple = ()                 # element map (element, type) pairs


def teacher_snapshots(teacher: NeuralNetworkPotential, n_structures: int = 64, n_snapshots: int = 2000,
                      every: int = 10, temperature: float = 300.0, min_molecules: int = 32,
                      max_molecules: int = 512, seed: int = 2023, md_precision: str = "mixed",
                      label_precision: str = "float64", tau_fs: float = 0.5, log=None):
    """Yield (structure, E*, F*) from teacher MD; F* = -grad E (input order, device FP64)."""
    from . import water
    from .simulation import BrendsenThermostat, MDSimulator, System

    rng = np.random.default_rng(seed)
    sizes = rng.integers(min_molecules, max_molecules + 1, size=n_structures)
    per = int(math.ceil(n_snapshots / n_structures))
    dt = 0.5 * units.FROM_FEMTO_SECOND
    count = 0
    for k, nm in enumerate(sizes):
        if count >= n_snapshots:
            break
        L = water.box_for_density(int(nm))
        s = Structure.from_dict(water.water_box(int(nm), L, seed=seed + 17 * (k + 1)))
        sysm = System.from_structure(s, teacher, temperature=temperature, seed=seed + k, force_mode="full",
                                     precision=md_precision, skin=1.0)
        md = MDSimulator(dt, thermostat=BrendsenThermostat(temperature, tau_fs * units.FROM_FEMTO_SECOND),
                         integrator="physical")
        for _ in range(per):
            if count >= n_snapshots:
                break
            md.run(sysm, every)
            snap = s.copy()
            snap.positions = sysm.positions.clone()
            e, f = teacher.energy_and_forces(snap, mode="full", precision=label_precision)
            count += 1
            yield snap, float(e), f.clone()
        if log:
            log(f"teacher structure {k}: {3 * int(nm)} atoms, {count} snapshots")


class StudentTrainer:
    """Adam training of a NeuralNetworkPotential's MLPs with the GPU kernels."""

    def __init__(self, student: NeuralNetworkPotential, precision: str = "mixed", w_energy: float = 1.0,
                 w_force: float = 10.0, lr: float = 1e-3, betas=(0.9, 0.999), eps: float = 1e-8,
                 weight_decay: float = 0.0, seed: int = 0):
        import torch

        self.pot = student
        self.dev = _lib.require_device()
        self.precision = precision
        self.g_prec, self.mlp_mode = PRECISION_MODES[precision]
        self.fdt = torch.float64 if self.g_prec == _lib.PT_F64 else torch.float32
        self.w_energy, self.w_force = float(w_energy), float(w_force)
        self.lr, self.betas, self.eps, self.wd = float(lr), betas, float(eps), float(weight_decay)
        self.rng = np.random.default_rng(seed)
        self.snapshots: List[Snapshot] = []
        self.rc = student.r_cutoff
        self.ld = max(ap.model_input_size for ap in student.atomic_potentials.values())
        self.step_count = 0
        self.mlps = None
        self.tables = {}
        self._raw: List[Tuple[Structure, float, object]] = []

    # ---- data ------------------------------------------------------------------
    def add(self, structure: Structure, energy: float, forces) -> None:
        """Queue a labelled structure (forces in input order, device FP64)."""
        self._raw.append((structure, float(energy), forces))

    def prepare(self, fit_scaler: bool = True) -> None:
        """Build neighbour lists and descriptors of all queued structures (and fit the scaler)."""
        import torch

        pot = self.pot
        for s, e, f in self._raw:
            emap = s.element_map
            nl = NeighborList(self.rc, emap.n_types).build(s.positions, s.atom_types, s.box_lengths,
                                                           s.type_counts(), with_rev=True)
            blocks = []
            for el in pot.elements:
                t = emap.element_to_atom_type.get(el)
                if t is None:
                    continue
                b, cnt = nl.type_block(t)
                if cnt:
                    blocks.append((el, b, cnt))
            G = torch.zeros((s.natoms, self.ld), dtype=self.fdt, device=self.dev)
            for el, b, cnt in blocks:
                tab = pot.atomic_potentials[el].descriptor.device_table(emap, self.dev)
                self.tables[(el, tuple(sorted(emap.element_to_atom_type.items())))] = tab
                acsf_forward(nl, tab, self.g_prec, cnt, center_begin=b, out=G[b:], ld=self.ld)
            F_int = gather_rows(nl.perm, f.to(torch.float64).contiguous())
            nl.release_workspace()
            self.snapshots.append(Snapshot(s.natoms, nl, blocks, G, F_int, e,
                                           tuple(sorted(emap.element_to_atom_type.items()))))
        self._raw = []
        if fit_scaler:
            from .descriptors import DescriptorScaler

            for el in pot.elements:
                params = None
                for sn in self.snapshots:
                    for e2, b, cnt in sn.blocks:
                        if e2 != el:
                            continue
                        n_sf = pot.atomic_potentials[el].model_input_size
                        x = sn.G[b:b + cnt, :n_sf]   # device rows: statistics on the GPU
                        params = DescriptorScaler.fit(x) if params is None else DescriptorScaler.partial_fit(params, x)
                pot.scalers_params[el] = params
            pot._mlp_key = None
        self._init_optimizer()

    def _init_optimizer(self) -> None:
        import torch

        self.mlps = self.pot._device_mlps(self.dev)
        self.grad = {e: torch.zeros(m.n_params, dtype=torch.float64, device=self.dev) for e, m in self.mlps.items()}
        self.m1 = {e: torch.zeros_like(g) for e, g in self.grad.items()}
        self.m2 = {e: torch.zeros_like(g) for e, g in self.grad.items()}
        nmax = max(sn.n for sn in self.snapshots)
        self.e_atom = torch.zeros(nmax, dtype=torch.float64, device=self.dev)
        self.dEdG = torch.zeros((nmax, self.ld), dtype=self.fdt, device=self.dev)
        self.T = torch.zeros((nmax, self.ld), dtype=self.fdt, device=self.dev)
        self.f_self = torch.zeros((nmax, 3), dtype=torch.float64, device=self.dev)
        self.F = torch.zeros((nmax, 3), dtype=torch.float64, device=self.dev)
        self.delta = torch.zeros((nmax, 3), dtype=torch.float64, device=self.dev)
        self.gy = torch.zeros(nmax, dtype=torch.float64, device=self.dev)
        self.loss = torch.zeros(2, dtype=torch.float64, device=self.dev)
        self.reducer = Reducer(nmax, self.dev)
        emax = max(sn.nl.n_pairs for sn in self.snapshots)
        self.edge = torch.zeros(3 * max(emax, 1), dtype=self.fdt, device=self.dev)

    # ---- one structure ------------------------------------------------------------
    def _table(self, sn: Snapshot, el: str):
        return self.tables[(el, sn.emap_key)]

    def _energy_forces(self, sn: Snapshot):
        """Student E (device scalar) and F (internal order) of a snapshot."""
        n = sn.n
        for el, b, cnt in sn.blocks:
            self.mlps[el].energy_grad(self.mlp_mode, cnt, sn.G[b:], self.ld, self.g_prec, self.e_atom[b:],
                                      self.dEdG[b:])
        E = self.reducer.sum(self.e_atom[:n])
        for el, b, cnt in sn.blocks:
            acsf_adjoint(sn.nl, self._table(sn, el), self.g_prec, cnt, self.dEdG[b:], self.ld, center_begin=b,
                         f_self=self.f_self[b:], edge_grad=self.edge)
        force_gather(sn.nl, self.edge, self.g_prec, out=self.F[:n])
        return E, self.F[:n]

    def _accumulate(self, sn: Snapshot, scale: float) -> None:
        n = sn.n
        E, F = self._energy_forces(sn)
        _lib.check(_lib.load().pt_train_seeds(n, F.data_ptr(), sn.F_ref.data_ptr(), E.data_ptr(), sn.E_ref,
                                              self.w_energy * scale, self.w_force * scale, self.delta.data_ptr(),
                                              self.gy.data_ptr(), self.loss.data_ptr(), _lib.stream_ptr()),
                   "pt_train_seeds")
        for el, b, cnt in sn.blocks:
            acsf_tangent(sn.nl, self._table(sn, el), self.g_prec, cnt, self.delta[:n], center_only=False,
                         center_begin=b, out=self.T[b:], ld=self.ld)
        for el, b, cnt in sn.blocks:
            self.mlps[el].train_grad(self.mlp_mode, cnt, sn.G[b:], self.T[b:], self.ld, self.g_prec, self.gy[b:],
                                     -1.0, self.grad[el])

    def gradient(self, indices: Sequence[int]):
        """Loss (energy term, force term) and its gradient over ``indices`` (batch mean)."""
        for g in self.grad.values():
            g.zero_()
        self.loss.zero_()
        scale = 1.0 / len(indices)
        for k in indices:
            self._accumulate(self.snapshots[k], scale)
        return self.loss.clone(), {e: g.clone() for e, g in self.grad.items()}

    def step(self, indices: Sequence[int]):
        loss, _ = self.gradient(indices)
        self.step_count += 1
        b1, b2 = self.betas
        for el, mlp in self.mlps.items():
            _lib.check(_lib.load().pt_adam_step(mlp.n_params, mlp.flat.data_ptr(), self.grad[el].data_ptr(),
                                                self.m1[el].data_ptr(), self.m2[el].data_ptr(), self.lr, b1, b2,
                                                self.eps, self.wd, self.step_count, _lib.stream_ptr()),
                       "pt_adam_step")
            mlp.repack()
        return loss

    def train(self, epochs: int, batch: int = 32, train_idx: Optional[Sequence[int]] = None, log=None,
              lr_schedule=None):
        """``epochs`` passes over ``train_idx`` in shuffled batches; ``lr_schedule(epoch)``
        (optional) sets the learning rate of each epoch."""
        import torch

        idx = np.arange(len(self.snapshots)) if train_idx is None else np.asarray(train_idx)
        history = []
        t0 = time.perf_counter()
        for ep in range(epochs):
            if lr_schedule is not None:
                self.lr = float(lr_schedule(ep))
            perm = self.rng.permutation(idx)
            tot = torch.zeros(2, dtype=torch.float64, device=self.dev)
            nb = 0
            for s0 in range(0, len(perm), batch):
                tot += self.step(perm[s0:s0 + batch])
                nb += 1
            l = (tot / max(nb, 1)).cpu().numpy()
            history.append((ep, float(l[0]), float(l[1])))
            if log:
                log(f"epoch {ep}: loss_E {l[0]:.4e} loss_F {l[1]:.4e} ({time.perf_counter() - t0:.1f} s)")
        self.sync_params()
        return history

    def sync_params(self) -> None:
        """Copy the device master parameters back into the potential's params dicts."""
        self.pot.models_params = {e: m.params_dict() for e, m in self.mlps.items()}
        self.pot._mlp_key = None
        self.mlps = self.pot._device_mlps(self.dev)
        # keep optimiser buffers aligned with the rebuilt device MLPs
        for e, m in self.mlps.items():
            if m.n_params != self.grad[e].numel():
                raise RuntimeError("parameter count changed")

    def evaluate(self, indices: Sequence[int]) -> Dict[str, float]:
        """Energy RMSE (meV/atom) and force RMSE (meV/Bohr) of the student over ``indices``."""
        import torch

        se, sf, nat = 0.0, 0.0, 0
        for k in indices:
            sn = self.snapshots[k]
            E, F = self._energy_forces(sn)
            se += ((float(E) - sn.E_ref) / sn.n) ** 2
            sf += float(((F - sn.F_ref) ** 2).sum())
            nat += sn.n
        mev = units.TO_MILLI_ELECTRON_VOLT
        return {"energy_rmse_mev_per_atom": math.sqrt(se / max(len(indices), 1)) * mev,
                "force_rmse_mev_per_bohr": math.sqrt(sf / max(3 * nat, 1)) * mev}


ASSET_DIR = __import__("pathlib").Path(__file__).resolve().parent / "assets" / "student_water"


def rank_teacher_g3(teacher: NeuralNetworkPotential, structures: Sequence[Structure], sample: int = 256,
                    seed: int = 0) -> Dict[str, Dict[str, List[Tuple[float, Tuple[float, float, float]]]]]:
    """The teacher's angular functions ranked by how strongly its forces depend on them.

    For central element e and function s: rms over atoms of dE_i/dG_is (the
    teacher's own gradient w.r.t. its raw descriptors) times rms over a sample of
    atoms of |dG_is/dr_i| (its self-Jacobian, ACSF.grad): the size of the force
    component carried by that function.  Returns, per element and neighbour pair
    ("H-O" ...), [(score, (eta, zeta, lambda)), ...] in decreasing score.
    """
    import torch

    from .bench_models import pair_key

    rng = np.random.default_rng(seed)
    acc: Dict[str, Dict[int, List[float]]] = {}
    for s in structures:
        ev = teacher._evaluation(s, "float64")
        ev.energy(teacher)
        for e, b, cnt in ev._blocks(teacher):
            acsf = teacher.atomic_potentials[e].descriptor
            n_sf = acsf.num_symmetry_functions
            w2 = (ev.dEdG[b:b + cnt, :n_sf].double() ** 2).sum(0).cpu().numpy()
            idx = s.select_host(e)
            pick = np.sort(rng.choice(idx, min(sample, idx.size), replace=False))
            J = acsf.grad(s, atom_index=pick, precision="float64")
            j2 = (J.double() ** 2).sum((0, 2)).cpu().numpy()
            d = acc.setdefault(e, {"w2": np.zeros(n_sf), "nw": 0, "j2": np.zeros(n_sf), "nj": 0})
            d["w2"] += w2
            d["nw"] += cnt
            d["j2"] += j2
            d["nj"] += pick.size
    out: Dict[str, Dict[str, List]] = {}
    for e, d in acc.items():
        score = np.sqrt(d["w2"] / max(d["nw"], 1)) * np.sqrt(d["j2"] / max(d["nj"], 1))
        acsf = teacher.atomic_potentials[e].descriptor
        nr = acsf.num_radial_symmetry_functions
        per: Dict[str, List] = {}
        for q, (sf, ne) in enumerate(acsf.angular_symmetry_functions):
            per.setdefault(pair_key(ne.neighbor_j, ne.neighbor_k), []).append(
                (float(score[nr + q]), (float(sf.eta), float(sf.zeta), float(sf.lambda0))))
        out[e] = {p: sorted(v, key=lambda t: -t[0]) for p, v in per.items()}
    return out


def select_g3(ranking, per_pair: int = 6) -> Dict[str, Dict[str, List[Tuple[float, float, float]]]]:
    """The ``per_pair`` best-ranked cells of every pair, in the teacher's column order."""
    from .bench_models import G3_GRID

    order = {c: q for q, c in enumerate(G3_GRID)}
    return {e: {p: sorted((c for _, c in v[:per_pair]), key=lambda c: order.get(tuple(c), 0)) for p, v in d.items()}
            for e, d in ranking.items()}


def cosine_schedule(lr: float, epochs: int, hold: float = 0.5, lr_min: float = 1e-5):
    """lr for the first ``hold`` of the epochs, then a cosine decay to lr_min."""
    h = int(hold * epochs)

    def f(ep: int) -> float:
        if ep < h:
            return lr
        t = (ep - h) / max(1, epochs - h - 1)
        return lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * min(1.0, t)))
    return f


def train_water_student(out_dir=ASSET_DIR, n_snapshots: int = 2000, n_structures: int = 64, every: int = 10,
                        epochs: int = 600, batch: int = 32, lr: float = 1e-3, w_force: float = 10.0,
                        seed: int = 2023, r_cutoff: float = 8.0, select: str = "sensitivity",
                        data=None, log=print) -> dict:
    """Generate teacher data, train the water student on the GPU and save it to ``out_dir``.

    Teacher: the 98,304-atom benchmark NNP (52 ACSF, 52-64-64-1); data: its own
    300 K MD on ``n_structures`` water boxes of 32..512 molecules at 1 g/cm^3,
    ``n_snapshots`` snapshots every ``every`` steps with FP64 labels; student:
    34 ACSF at cutoff ``r_cutoff`` whose 6 G3 per pair are the teacher's most
    force-relevant ones (``select="sensitivity"``, rank_teacher_g3 on the first
    snapshots) or its zeta = 1 ones (``"zeta1"``), MLP 34-16-16-1, mixed
    precision, Adam (lr for half the epochs, then a cosine decay), loss
    E-MSE/atom + w_force F-MSE.  ``data``: labelled (structure, E, F) triples
    to reuse instead of generating them.
    """
    from pathlib import Path

    from . import bench_models

    t0 = time.perf_counter()
    _, teacher = bench_models.water_teacher(precision="mixed", seed=seed)
    log(f"[student] teacher ready ({time.perf_counter() - t0:.1f} s)")
    if data is None:
        data = list(teacher_snapshots(teacher, n_structures=n_structures, n_snapshots=n_snapshots, every=every,
                                      seed=seed, log=log))
    log(f"[student] {len(data)} snapshots labelled ({time.perf_counter() - t0:.1f} s)")
    if select == "sensitivity":
        ranking = rank_teacher_g3(teacher, [sn for sn, _, _ in data[:: max(1, len(data) // 48)]], seed=seed)
        spec = bench_models.student_spec(r_cutoff, select_g3(ranking))
        spec["ranking"] = {e: {p: [[sc, list(c)] for sc, c in v] for p, v in d.items()} for e, d in ranking.items()}
    else:
        spec = bench_models.student_spec(r_cutoff)
    student = bench_models.student_from_spec(spec, seed=seed + 5)
    tr = StudentTrainer(student, precision="mixed", w_energy=1.0, w_force=w_force, lr=lr, seed=seed)
    for snap, e, f in data:
        tr.add(snap, e, f)
    tr.prepare(fit_scaler=True)
    n = len(tr.snapshots)
    perm = np.random.default_rng(seed).permutation(n)
    n_val = max(1, n // 10)
    val, train = np.sort(perm[:n_val]), np.sort(perm[n_val:])
    before = tr.evaluate(val)
    log(f"[student] untrained: {before}")
    t1 = time.perf_counter()
    hist = tr.train(epochs, batch=batch, train_idx=train, lr_schedule=cosine_schedule(lr, epochs),
                    log=lambda m: log(f"[student] {m}"))
    t_train = time.perf_counter() - t1
    after_val = tr.evaluate(val)
    after_train = tr.evaluate(train)
    log(f"[student] trained: validation {after_val}, training {after_train} ({t_train:.1f} s)")
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    student.save(out)
    (out / bench_models.STUDENT_SPEC).write_text(json.dumps(spec, indent=1))
    f_rms = math.sqrt(sum(float((sn.F_ref ** 2).sum()) for sn in tr.snapshots) /
                      (3 * sum(sn.n for sn in tr.snapshots))) * units.TO_MILLI_ELECTRON_VOLT
    metrics = {"snapshots": n, "validation": int(n_val), "epochs": epochs, "batch": batch, "lr": lr,
               "lr_schedule": f"{lr} for {epochs // 2} epochs, then cosine to 1e-5",
               "w_force": w_force, "r_cutoff": r_cutoff, "select": select,
               "adam_steps": tr.step_count, "train_seconds": t_train,
               "untrained_validation": before, "validation_metrics": after_val, "training_metrics": after_train,
               "teacher_force_rms_mev_per_bohr": f_rms,
               "loss_history": [(int(a), float(b), float(c)) for a, b, c in hist[:: max(1, len(hist) // 60)]]}
    (out / "metrics.json").write_text(json.dumps(metrics, indent=1))
    return metrics


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="train the water student NNP on teacher MD data (GPU)")
    ap.add_argument("--out", default=str(ASSET_DIR))
    ap.add_argument("--snapshots", type=int, default=2000)
    ap.add_argument("--structures", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=600)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--rc", type=float, default=8.0, help="student cutoff (Bohr)")
    ap.add_argument("--select", default="sensitivity", choices=("sensitivity", "zeta1"))
    a = ap.parse_args()
    train_water_student(a.out, a.snapshots, a.structures, epochs=a.epochs, batch=a.batch, lr=a.lr,
                        r_cutoff=a.rc, select=a.select, log=lambda m: print(m, flush=True))
