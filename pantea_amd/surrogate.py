"""Surrogate-driven MD with a judge and fallback.

A student NNP drives velocity-Verlet MD.  Every ``judge_every`` steps the
teacher NNP (FP64 by default) is evaluated on the current positions and its
forces are compared with the student's (RMSE over all force components, or
over the atoms of a random sample when ``judge_atoms`` is set).  When the
RMSE exceeds ``threshold`` the run rewinds to the last state that passed and
the teacher drives the next ``fallback_steps`` steps; then the student takes
over again from the state the teacher reached.  A student is on probation
until it has passed a judge, and again after every failure: it is then
judged every ``probation_every`` steps (default ``judge_every``), so a
student that fails wastes fewer discarded steps.

The fallback teacher may run at a lower precision than the judge (e.g. the
mixed-precision teacher under an FP64 judge).  It is then judged like the
student: at every ``judge_every`` steps and at the end of its window, and a
failure rewinds to the last state that passed and hands the window to the
judge's own model, which needs no judging.  A fallback at the judge's
precision is the judge's model and is never judged.

A pre-judge (``prejudge_margin``) watches the student's own inputs: after
every student step the device counts descriptor rows outside the student's
scaler range widened by the margin (pt_scaler_extrapolation, the check of
descriptors/scaler.py:286-295), and every ``prejudge_every`` steps the host
reads that count; rows out of range call the judge at once instead of at the
next judge interval, so a student that leaves its training distribution is
judged before it has driven long on extrapolated inputs.

A sampled judge evaluates the teacher only where the sampled atoms' full
forces need it: F_k = -sum_i dE_i/dr_k involves the centres i within the
cutoff of k, so descriptors, MLP and adjoint run on the closed
neighbourhoods of the sample (pt_nbr_mark_rows), and the force gather then
gives exact teacher forces on the sampled atoms.  Its cost is the share of
atoms in those neighbourhoods instead of the whole system.

Student, judge and fallback potentials share one device state.  The
neighbour list has the driver's cutoff (plus the skin), so a student with a
shorter cutoff than the teacher stages only its own neighbours; a judge of
such a student builds the teacher's list for the current positions, and a
change of driver rebuilds the list at the new driver's cutoff.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib
from .simulation import BrendsenThermostat, System
from .units import units


@dataclass
class SurrogateStats:
    steps: int = 0                 # net trajectory length (steps kept)
    student_steps: int = 0         # steps driven by the student (kept or discarded)
    fallback_steps: int = 0        # kept steps driven by the teacher (either precision)
    reference_steps: int = 0       # of those, steps driven by the judge's own model
    discarded_steps: int = 0       # steps thrown away by rewinds
    judges: int = 0
    failures: int = 0              # failed judges (student or judged fallback)
    fallback_failures: int = 0     # of those, failures of the judged fallback
    judge_rmse: list = field(default_factory=list)   # Ha/Bohr, one per judge
    judge_driver: list = field(default_factory=list)  # "student" or "fallback", one per judge
    list_overflows: int = 0        # judged windows redone because a gated list build overflowed
    prejudge_triggers: int = 0     # judges called early by the extrapolation count
    extrapolated_rows: int = 0     # student descriptor rows seen out of range (pre-judge)
    # kept trajectory as windows (first step, end step, driver, judge RMSE in Ha/Bohr; None for
    # steps of the judge's own model, which needs no judging)
    windows: list = field(default_factory=list)
    seconds: float = 0.0

    @property
    def fallback_rate(self) -> float:
        """Share of judged windows that failed."""
        return self.failures / self.judges if self.judges else 0.0

    @property
    def teacher_share(self) -> float:
        """Share of the kept trajectory driven by the teacher."""
        return self.fallback_steps / self.steps if self.steps else 0.0

    def as_dict(self) -> dict:
        mev = units.TO_MILLI_ELECTRON_VOLT
        r = np.asarray(self.judge_rmse, dtype=np.float64) * mev
        return {"steps": self.steps, "student_steps": self.student_steps, "fallback_steps": self.fallback_steps,
                "reference_steps": self.reference_steps,
                "discarded_steps": self.discarded_steps, "judges": self.judges, "failures": self.failures,
                "fallback_failures": self.fallback_failures, "list_overflows": self.list_overflows,
                "prejudge_triggers": self.prejudge_triggers, "extrapolated_rows": self.extrapolated_rows,
                "student_kept_steps": sum(b - a for a, b, d, _ in self.windows if d == "student"),
                "fallback_rate": self.fallback_rate, "teacher_share": self.teacher_share,
                "judge_rmse_mev_per_bohr_median": float(np.median(r)) if r.size else None,
                "judge_rmse_mev_per_bohr_max": float(np.max(r)) if r.size else None,
                "seconds": self.seconds}


class SurrogateMD:
    """Student-driven MD judged by the teacher (see the module docstring)."""

    def __init__(self, teacher, student, structure, temperature: float = 300.0, seed: int = 2023,
                 time_step: float = 0.5 * units.FROM_FEMTO_SECOND, integrator: str = "physical",
                 student_precision: str = "mixed", judge_precision: str = "float64",
                 fallback_precision: str = "float64", judge_every: int = 20,
                 threshold: float = 5.0 / units.TO_MILLI_ELECTRON_VOLT, fallback_steps: int = 200,
                 skin: float = 1.0, thermostat: Optional[BrendsenThermostat] = None,
                 judge_atoms: Optional[int] = None, judge_seed: int = 0,
                 probation_every: Optional[int] = None, graph: bool = False,
                 prejudge_margin: Optional[float] = None, prejudge_every: int = 5):
        import torch

        if integrator not in ("reference", "physical"):
            raise ValueError(f"unknown integrator '{integrator}'")
        if judge_every < 1 or fallback_steps < 1:
            raise ValueError("judge_every and fallback_steps must be >= 1")
        if probation_every is not None and not 1 <= probation_every <= judge_every:
            raise ValueError("probation_every must be in [1, judge_every]")
        self.system = System.from_structure(structure, student, temperature=temperature, seed=seed,
                                            force_mode="full", precision=student_precision, skin=skin,
                                            graph=graph)
        md = self.system._engine()
        self.md = md
        self.f_student = md.ff
        self.f_judge = md.add_field(teacher, judge_precision)
        self.f_fallback = md.add_field(teacher, fallback_precision)
        self.f_student.label = f"student-{student_precision}"
        self.f_judge.label = f"teacher-{judge_precision}"
        self.f_fallback.label = f"teacher-{fallback_precision}"
        # a fallback below the judge's precision is judged; the judge's own model is not
        self.judged_fallback = fallback_precision != judge_precision
        # the list may have grown: forces of the driver at the current positions
        md.compute(md.force, "full")
        self.dt = float(time_step)
        self.integrator = integrator
        self.thermostat = thermostat
        self.judge_every = int(judge_every)
        self.probation_every = int(probation_every) if probation_every is not None else self.judge_every
        self.on_probation = True   # until the student passes its first judge
        self.threshold = float(threshold)
        self.fallback_len = int(fallback_steps)
        self.f_judge_out = torch.zeros_like(md.force)
        self.judge_atoms = judge_atoms
        self.rng = np.random.default_rng(judge_seed)
        self._flags = None
        self.stats = SurrogateStats()
        self.prejudge_margin = prejudge_margin
        self.prejudge_every = int(prejudge_every)
        if prejudge_margin is not None:
            from .descriptors import params_to_device

            self._extra = torch.zeros(3, dtype=torch.int64, device=md.dev)
            self._extra_params = {e: params_to_device(student.scalers_params[e], md.dev)
                                  for e in student.elements if student.scalers_params.get(e) is not None}
        self.mode = "student"
        self.fallback_left = 0
        self.good = md.save_state()
        self.good_step = 0

    # ---- judge -------------------------------------------------------------------
    def judge(self) -> float:
        """Force RMSE (Ha/Bohr) of the driving student against the judge at the current positions."""
        md = self.md
        if self.judge_atoms and self.judge_atoms < md.n:
            sel = np.sort(self.rng.choice(md.n, int(self.judge_atoms), replace=False)).astype(np.int32)
            f = self.judge_forces(sel)
            diff = md.force[self.sel_dev] - f
        else:
            md.compute(self.f_judge_out, "full", field=self.f_judge)
            diff = md.force - self.f_judge_out
        return math.sqrt(float((diff * diff).mean()))

    def judge_forces(self, sel: np.ndarray):
        """Judge forces -grad E on the atoms ``sel`` (internal order), evaluating only their neighbourhoods."""
        import torch

        from .engine import acsf_adjoint, acsf_forward, force_gather

        md = self.md
        ff = self.f_judge
        nl = md.list_for(ff)   # the teacher's own list when its cutoff exceeds the student's
        dev = md.dev
        lib = _lib.load()
        if self._flags is None or self._flags.numel() != md.n:
            self._flags = torch.zeros(md.n, dtype=torch.uint8, device=dev)
        self._flags.zero_()
        self.sel_dev = torch.as_tensor(np.ascontiguousarray(sel, dtype=np.int32), device=dev)
        sel_nl = self.sel_dev
        if nl is not md.nl:
            # the sampled atoms in the auxiliary list's order: inverse of its permutation
            inv = torch.empty(md.n, dtype=torch.int32, device=dev)
            inv[nl.perm.long()] = torch.arange(md.n, dtype=torch.int32, device=dev)
            sel_nl = inv[self.sel_dev.long()].contiguous()
        _lib.check(lib.pt_nbr_mark_rows(int(sel.size), sel_nl.data_ptr(), nl.row_off.data_ptr(),
                                        nl.nbr.data_ptr(), self._flags.data_ptr(), _lib.stream_ptr()),
                   "pt_nbr_mark_rows")
        cent = np.flatnonzero(self._flags.cpu().numpy()).astype(np.int32)
        edt = torch.float64 if ff.g_prec == _lib.PT_F64 else torch.float32
        if ff.edge is None or ff.edge.numel() < nl.n_pairs * 3 or ff.edge.dtype != edt:
            ff.edge = torch.zeros(int(nl.n_pairs * 3 * 1.2) + 3, dtype=edt, device=dev)
        if not ff.covered:
            ff.edge.zero_()
        fdt = ff.G.dtype
        blocks = []
        for e, b, cnt in ff._blocks(nl):
            c = cent[(cent >= b) & (cent < b + cnt)]
            if c.size:
                blocks.append((e, torch.as_tensor(c, device=dev)))
        m = sum(int(c.numel()) for _, c in blocks)
        G = torch.zeros((m, ff.ld), dtype=fdt, device=dev)
        dEdG = torch.zeros_like(G)
        e_atom = torch.zeros(m, dtype=torch.float64, device=dev)
        o = 0
        for e, c in blocks:
            k = int(c.numel())
            acsf_forward(nl, ff.tables[e], ff.g_prec, k, centers=c, out=G[o:], ld=ff.ld)
            ff.mlps[e].energy_grad(ff.mlp_mode, k, G[o:], ff.ld, ff.g_prec, e_atom[o:], dEdG[o:])
            acsf_adjoint(nl, ff.tables[e], ff.g_prec, k, dEdG[o:], ff.ld, centers=c, edge_grad=ff.edge)
            o += k
        self.judge_centres = m
        force_gather(nl, ff.edge, ff.g_prec, out=self.f_judge_out)
        return self.f_judge_out[sel_nl.long()]

    # ---- driving -----------------------------------------------------------------
    def _to_student(self) -> None:
        md = self.md
        if self.mode == "reference" and self.stats.steps > self.good_step:
            self.stats.windows.append((self.good_step, self.stats.steps, "reference", None))
        md.use(self.f_student)
        md.compute(md.force, "full")
        self.good = md.save_state()
        self.good_step = self.stats.steps
        self.mode = "student"

    def _to_fallback(self, reference: bool = False) -> None:
        """Rewind to the last state that passed and hand the next window to the
        fallback teacher (or, ``reference``, to the judge's own model)."""
        md = self.md
        discarded = self.stats.steps - self.good_step
        self.stats.discarded_steps += discarded
        if self.mode == "fallback":   # the fallback's own steps since its last pass
            self.stats.fallback_steps -= discarded
        md.restore_state(self.good)
        self.stats.steps = self.good_step
        reference = reference or not self.judged_fallback
        md.use(self.f_judge if reference else self.f_fallback)
        md.compute(md.force, "full")
        self.mode = "reference" if reference else "fallback"
        self.fallback_left = self.fallback_len

    def _judge_step(self) -> bool:
        """Judge the driver at the current positions; on a pass the state becomes the rewind point."""
        st = self.stats
        rmse = self.judge()
        if self.prejudge_margin is not None:
            self._prejudge_due()   # a judge covers every step before it
        # the judge synchronises: check the gated list builds of the window with it
        if not self.md.lists_ok():
            self.md.grow_lists()
            st.list_overflows += 1
            rmse = math.inf
        st.judges += 1
        st.judge_rmse.append(rmse)
        st.judge_driver.append(self.mode)
        if rmse <= self.threshold:
            st.windows.append((self.good_step, st.steps, self.mode, rmse))
            self.good = self.md.save_state()
            self.good_step = st.steps
            return True
        st.failures += 1
        if self.mode == "fallback":
            st.fallback_failures += 1
        return False

    def run(self, n_steps: int, judge_tail: bool = False) -> SurrogateStats:
        """Advance the kept trajectory by ``n_steps`` steps.  judge_tail: the steps after the
        last judge are judged too (and redone by the fallback if they fail), so that every kept
        step lies in a judged window (or one of the judge's own model)."""
        import torch

        md = self.md
        st = self.stats
        target = st.steps + int(n_steps)
        t0 = time.perf_counter()
        while True:
            self._advance(target)
            # steps after the last judge: their list builds are checked before they count
            if not md.lists_ok():
                md.grow_lists()
                st.list_overflows += 1
                self._to_fallback()
                continue
            if judge_tail and st.steps > self.good_step:
                if self.mode == "reference":
                    st.windows.append((self.good_step, st.steps, "reference", None))
                    self.good = md.save_state()
                    self.good_step = st.steps
                elif not self._judge_step():
                    self._to_fallback(reference=self.mode == "fallback")
                    continue
            break
        torch.cuda.synchronize()
        st.seconds += time.perf_counter() - t0
        return st

    def _count_extrapolation(self) -> None:
        """Add the student's out-of-range descriptor rows of this step to the device counter."""
        from .descriptors import extrapolation_counts

        ff = self.f_student
        nl = self.md.nl
        for e, b, cnt in ff._blocks(nl):
            p = self._extra_params.get(e)
            if p is not None:
                n_sf = ff.pot.atomic_potentials[e].model_input_size
                extrapolation_counts(ff.G[b:b + cnt, :n_sf], p, self._extra, self.prejudge_margin)

    def _prejudge_due(self) -> bool:
        """Whether the student extrapolated since the last read (a host read)."""
        rows = int(self._extra[0].item())
        if rows:
            self.stats.extrapolated_rows += rows
            self._extra.zero_()
        return rows > 0

    def _advance(self, target: int) -> None:
        md = self.md
        st = self.stats
        while st.steps < target:
            md.step(self.dt, self.integrator, "full", self.thermostat)
            st.steps += 1
            if self.mode == "student":
                st.student_steps += 1
                every = self.probation_every if self.on_probation else self.judge_every
                early = False
                if self.prejudge_margin is not None:
                    self._count_extrapolation()
                    if (st.steps - self.good_step) % self.prejudge_every == 0 and \
                            st.steps - self.good_step < every and self._prejudge_due():
                        early = True
                        st.prejudge_triggers += 1
                if early or st.steps - self.good_step >= every:
                    passed = self._judge_step()
                    self.on_probation = not passed
                    if not passed:
                        self._to_fallback()
                continue
            st.fallback_steps += 1
            if self.mode == "reference":
                st.reference_steps += 1
                self.fallback_left -= 1
                if self.fallback_left == 0:
                    self._to_student()
                continue
            # judged fallback: at the judge interval and before handing back to the student
            self.fallback_left -= 1
            due = st.steps - self.good_step >= self.judge_every or self.fallback_left == 0
            if due and not self._judge_step():
                self._to_fallback(reference=True)
            elif self.fallback_left == 0:
                self._to_student()

    @property
    def positions(self):
        return self.md.positions()
