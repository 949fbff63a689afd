"""Molecular dynamics on the GPU (simulation/molecular_dynamics.py:16-90,
simulation/system.py:55-164, simulation/thermostat.py:12-66,
simulation/simulate.py:31-88).

The state of a System lives on the device in the neighbour list's internal
order (atoms grouped by type and cell).  One MD step is

  pt_verlet_positions  x <- wrap(x + v dt + 1/2 a dt^2)
  [pt_nbr_build         when the list is stale]
  descriptors -> MLP -> adjoint -> force gather   (the potential's kernels)
  pt_verlet_velocities v <- v + 1/2 (a + a') dt
  [pt_kinetic_energy + pt_berendsen_rescale]

Two integrator modes: ``"reference"`` (a = F, the reference's update that
never divides by the mass, molecular_dynamics.py:16-30) and ``"physical"``
(a = F / m).  Two force modes: ``"reference"`` (the reference's self term)
and ``"full"`` (-grad E).

Neighbour lists are built with cutoff rc + skin and rebuilt as soon as the
two largest atom displacements since the last build sum to more than the
skin (no pair can have moved by more); the descriptor kernels re-apply the
exact mask (r <= rc, r > 0), so every step sees exactly the reference's
neighbour set.  In a periodic box that decision is taken on the device
(pt_skin_check; without a skin every step rebuilds) and the rebuild runs
behind its flag (pt_nbr_build_async, pt_permute_state_gated), so a step
never waits for the GPU; the list's totals are verified whenever the host
reads the state.
"""

from __future__ import annotations

import ctypes
from pathlib import Path
from typing import Optional

import numpy as np

from . import _lib
from .atoms import ElementMap, host_box
from .engine import (NeighborList, Reducer, acsf_adjoint, acsf_forward, force_gather, force_gather_transposed,
                     gather_rows, scatter_rows)
from .potential import PRECISION_MODES, NeuralNetworkPotential
from .units import units

KB = units.BOLTZMANN_CONSTANT


class BrendsenThermostat:
    """Berendsen velocity rescaling (thermostat.py:12-66)."""

    def __init__(self, target_temperature: float, time_constant: float):
        self.target_temperature = float(target_temperature)
        self.time_constant = float(time_constant)


class _ForceField:
    """Descriptor/MLP buffers and launch sequence of one potential at one precision."""

    def __init__(self, pot: NeuralNetworkPotential, precision: str, emap, n: int, dev):
        import torch

        self.pot = pot
        self.precision = precision
        self.label = precision  # timer prefix; owners may rename it
        self.g_prec, self.mlp_mode = PRECISION_MODES[precision]
        self.emap = emap
        self.n = n
        self.dev = dev
        self.rc = pot.r_cutoff
        self.ld = max(ap.model_input_size for ap in pot.atomic_potentials.values())
        fdt = torch.float64 if self.g_prec == _lib.PT_F64 else torch.float32
        self.G = torch.zeros((n, self.ld), dtype=fdt, device=dev)
        self.dEdG = torch.zeros((n, self.ld), dtype=fdt, device=dev)
        self.e_atom = torch.zeros(n, dtype=torch.float64, device=dev)
        self.f_self = torch.zeros((n, 3), dtype=torch.float64, device=dev)
        self.edge = None
        self.reducer = Reducer(n, dev)
        self.tables = {e: pot.atomic_potentials[e].descriptor.device_table(emap, dev)
                       for e in pot.elements if e in emap.element_to_atom_type}
        # descriptor tables hold no centre element: elements whose tables are
        # byte-identical share one launch (one grid tail instead of several)
        self._table_key = {e: bytes(t.host) for e, t in self.tables.items()}
        self.covered = all(e in pot.atomic_potentials for e in emap.unique_elements)
        self.mlps = pot._device_mlps(dev)
        self.energy = None

    def refresh(self) -> None:
        """Pick up changed model parameters of the potential."""
        self.mlps = self.pot._device_mlps(self.dev)

    def _blocks(self, nl: NeighborList):
        for e in self.pot.elements:
            t = self.emap.element_to_atom_type.get(e)
            if t is None:
                continue
            b, cnt = nl.type_block(t)
            if cnt > 0:
                yield e, b, cnt

    def _desc_runs(self, nl: NeighborList):
        """(table, first centre, centres) of each descriptor launch: adjacent element
        blocks (internal order is type-sorted) with identical tables form one run."""
        runs = []
        for e, b, cnt in sorted(self._blocks(nl), key=lambda x: x[1]):
            if runs and runs[-1][0] == self._table_key[e] and runs[-1][2] + runs[-1][3] == b:
                runs[-1][3] += cnt
            else:
                runs.append([self._table_key[e], e, b, cnt])
        return [(self.tables[e], b, cnt) for _, e, b, cnt in runs]

    def compute(self, nl: NeighborList, out, force_mode: str, with_energy: bool = False, mark=None):
        import torch

        mark = mark or (lambda phase, start: None)
        blocks = list(self._blocks(nl))
        runs = self._desc_runs(nl)
        mark("acsf_forward", True)
        for tab, b, cnt in runs:
            acsf_forward(nl, tab, self.g_prec, cnt, center_begin=b, out=self.G[b:], ld=self.ld)
        mark("acsf_forward", False)
        mark("mlp", True)
        for e, b, cnt in blocks:
            self.mlps[e].energy_grad(self.mlp_mode, cnt, self.G[b:], self.ld, self.g_prec, self.e_atom[b:],
                                     self.dEdG[b:])
        mark("mlp", False)
        if with_energy:
            self.energy = self.reducer.sum(self.e_atom)
        full = force_mode == "full"
        if full:
            need = nl.n_pairs * 3
            edt = torch.float64 if self.g_prec == _lib.PT_F64 else torch.float32
            if self.edge is None or self.edge.numel() < need:
                self.edge = torch.zeros(int(need * 1.2) + 3, dtype=edt, device=self.dev)
            if not self.covered:
                self.edge.zero_()
        target = self.f_self if full else out
        if not self.covered:
            target.zero_()
        mark("acsf_adjoint", True)
        for tab, b, cnt in runs:
            # -grad E: edge gradients in the reverse edges' slots, so the gather reads rows contiguously
            acsf_adjoint(nl, tab, self.g_prec, cnt, self.dEdG[b:], self.ld, center_begin=b,
                         f_self=target[b:], edge_grad=self.edge if full else None, transposed=full)
        mark("acsf_adjoint", False)
        if full:
            mark("force_gather", True)
            force_gather_transposed(nl, self.edge, self.g_prec, self.f_self, out=out)
            mark("force_gather", False)
        return out


class _DeviceMD:
    """Device-resident MD state of one System (internal atom order).

    ``extra`` potentials (e.g. a surrogate's judge) share the state.  The
    neighbour list follows the driving potential's cutoff (plus the skin);
    a potential with a longer cutoff is evaluated on an auxiliary list built
    for the current positions when it is asked for (``list_for``), and the
    list switches cutoff when the driver does (``use``).

    Periodic boxes run the step without any host read ("gated" step): the
    skin check sets a device flag, and the rebuild (pt_nbr_build_async) and
    the permutation of the state into the new order are launched every step
    behind that flag, so the host never waits for the GPU.  The list's
    totals stay on the device; they are checked when the host synchronises
    anyway (every ``check_every`` steps of run(), a judge, an energy read).
    The descriptor kernels are launched with a row bound (the longest row of
    the last synchronous build plus a margin) and the list arrays are sized
    with a margin over its pair count: a step that exceeded either is found
    at the next check, and run() then redoes the steps since its checkpoint
    with larger bounds.  Structures without a box use the synchronous step.
    """

    ROW_MARGIN = 16          # row bound = longest row + this
    PAIR_MARGIN = 1.3        # list arrays = pairs * this

    def __init__(self, system: "System", skin: float, extra=(), check_every: int = 50, sync_free: bool = True):
        import torch

        pot = system.potential
        s = system.structure
        self.dev = s.device
        self.n = s.natoms
        self.emap = s.element_map
        self.box = s.box_lengths
        self.skin = float(skin)
        self.gated = self.box is not None and sync_free
        self.check_every = int(check_every)
        n = self.n
        self.fields = {(id(pot), system.precision): _ForceField(pot, system.precision, self.emap, n, self.dev)}
        for p, prec in extra:
            self.fields.setdefault((id(p), prec), _ForceField(p, prec, self.emap, n, self.dev))
        self.ff = self.fields[(id(pot), system.precision)]
        self.rc = self.ff.rc          # the list's cutoff, without the skin: the driver's
        self._aux = {}                # cutoff -> auxiliary list of a longer-ranged potential
        self.type_counts = s.type_counts()
        self.timers = None  # {phase: [(start_event, end_event), ...]} when profiling
        self._builds = 0
        self.redos = 0        # stretches redone after a list overflow
        self.use_graph = False
        self._graph = None
        self._graph_key = None
        self._since = 0       # gated steps since the last verified checkpoint
        self._ck = None
        self._args = None
        self.flag = torch.ones(1, dtype=torch.int32, device=self.dev)
        self.totals = torch.zeros(4, dtype=torch.int64, device=self.dev)
        self.disp = torch.zeros(514, dtype=torch.float64, device=self.dev)
        self.reducer = Reducer(n, self.dev)
        # state in input order -> internal order of the first build
        self._nl = NeighborList(self.rc + self.skin, self.emap.n_types)
        nl = self._nl.build(s.positions, s.atom_types, self.box, self.type_counts, with_rev=True)
        self._builds += 1
        perm = nl.perm
        self.orig = perm.clone()                                     # internal -> input index
        self.types = gather_rows(perm, s.atom_types.to(torch.int32).reshape(-1, 1)).reshape(-1).contiguous()
        self.vel = gather_rows(perm, system.velocities.contiguous())
        mass = system.masses.reshape(-1, 1).contiguous()
        self.mass = gather_rows(perm, mass).reshape(-1).contiguous()
        self.inv_mass = (1.0 / self.mass).contiguous()
        self.ref_pos = nl.pos4.clone()
        self.force = torch.zeros((n, 3), dtype=torch.float64, device=self.dev)
        self.force_new = torch.zeros_like(self.force)
        self._size_bounds()

    # ---- list bounds of the gated step ------------------------------------------
    def _size_bounds(self, grow: bool = False) -> None:
        """Row bound and list capacity from the last synchronous build (grown after an overflow)."""
        import torch

        nl = self._nl
        self.pos_in = torch.empty_like(nl.pos4)
        self.tmp3 = torch.empty((self.n, 3), dtype=torch.float64, device=self.dev)
        self.tmp1 = torch.empty(self.n, dtype=torch.float64, device=self.dev)
        self.tmpi = torch.empty(self.n, dtype=torch.int32, device=self.dev)
        # the gated rebuild's permutation of v, F, m, 1/m, input index, type: 72 bytes per atom
        self.tmp_state = torch.empty(self.n * 72, dtype=torch.uint8, device=self.dev)
        f = 2.0 if grow else 1.0
        self.row_bound = int(f * nl.max_row) + self.ROW_MARGIN
        cap = int(f * self.PAIR_MARGIN * nl.n_pairs) + 1024
        if cap > nl.cap:
            nl.cap = cap
            nl.nbr = torch.empty(cap, dtype=torch.int32, device=self.dev)
            nl.rev = torch.empty(cap, dtype=torch.int32, device=self.dev)
            # the list content was lost with the arrays: build it again
            self._rebuild_sync(nl.pos4.clone())
        self.totals.zero_()

    def add_field(self, pot, precision: str) -> _ForceField:
        """Evaluate ``pot`` at ``precision`` on this state too."""
        key = (id(pot), precision)
        if key not in self.fields:
            self.fields[key] = _ForceField(pot, precision, self.emap, self.n, self.dev)
        return self.fields[key]

    # the driving potential's buffers (kept for callers that read them)
    @property
    def pot(self):
        return self.ff.pot

    @property
    def energy(self):
        return self.ff.energy

    @property
    def e_atom(self):
        return self.ff.e_atom

    def field(self, pot, precision: str) -> _ForceField:
        return self.fields[(id(pot), precision)]

    def use(self, field: _ForceField) -> None:
        """Make ``field`` the driving potential of step(); the list takes its cutoff
        (a synchronous rebuild when the cutoff changes)."""
        if field.rc != self.rc:
            self.verify()
            self.rc = field.rc
            pos = self._nl.pos4.clone()
            self._nl = NeighborList(self.rc + self.skin, self.emap.n_types)
            self._rebuild_sync(pos)
            self._size_bounds()
        self.ff = field

    def list_for(self, field: _ForceField) -> NeighborList:
        """The list to evaluate ``field`` on at the current positions: the MD's own list
        when the field's cutoff is within the driver's, else an auxiliary list built now
        (synchronously; its ``perm`` maps its order to the MD's internal order)."""
        if field.rc <= self.rc:
            return self._nl
        aux = self._aux.get(field.rc)
        if aux is None:
            aux = self._aux[field.rc] = NeighborList(field.rc, self.emap.n_types)
        aux.build(self._nl.pos4, self.types, self.box, self.type_counts, with_rev=True)
        return aux

    def _mark(self, phase: str, start: bool) -> None:
        if self.timers is None:
            return
        import torch

        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        if start:
            self.timers.setdefault(phase, []).append([ev, None])
        else:
            self.timers[phase][-1][1] = ev

    @property
    def nl(self) -> NeighborList:
        return self._nl

    @property
    def n_builds(self) -> int:
        """Neighbour builds so far (reads the device counter of the gated steps: a sync)."""
        return self._builds + (int(self.totals[3].item()) if self.gated else 0)

    # ---- neighbour maintenance ---------------------------------------------
    def _permute_state(self, p) -> None:
        self.orig = gather_rows(p, self.orig.reshape(-1, 1)).reshape(-1).contiguous()
        self.types = gather_rows(p, self.types.reshape(-1, 1)).reshape(-1).contiguous()
        self.vel = gather_rows(p, self.vel)
        self.force = gather_rows(p, self.force)
        self.mass = gather_rows(p, self.mass.reshape(-1, 1)).reshape(-1).contiguous()
        self.inv_mass = gather_rows(p, self.inv_mass.reshape(-1, 1)).reshape(-1).contiguous()

    def _rebuild_sync(self, positions=None) -> None:
        """New list and internal order from ``positions`` (current internal order, n x 3 or 4),
        read back synchronously."""
        nl = self._nl
        src = nl.pos4.clone() if positions is None else positions
        nl.build(src, self.types, self.box, self.type_counts, with_rev=True)
        self._permute_state(nl.perm)
        self.ref_pos.copy_(nl.pos4)
        self._builds += 1
        self.row_bound = max(getattr(self, "row_bound", 0), nl.max_row)

    def _rebuild_gated(self) -> None:
        """Skin check, then (behind its device flag) rebuild and reorder the state: no host read."""
        lib = _lib.load()
        nl = self._nl
        st = _lib.stream_ptr()
        box = host_box(self.box)
        box_p = ctypes.cast(box, ctypes.c_void_p)
        g = self.flag.data_ptr()
        if self.skin > 0.0:
            _lib.check(lib.pt_skin_check(self.n, nl.pos4.data_ptr(), self.ref_pos.data_ptr(), box_p, self.skin,
                                         self.disp.data_ptr(), g, st), "pt_skin_check")
        # without a skin the flag keeps the 1 it was created with: every step rebuilds, and
        # neither the displacement check nor the reference positions are needed
        _lib.check(lib.pt_copy_gated(4 * self.n, 8, nl.pos4.data_ptr(), self.pos_in.data_ptr(), g, st),
                   "pt_copy_gated")
        _lib.check(lib.pt_nbr_build_async(self.n, self.pos_in.data_ptr(), 4, self.types.data_ptr(), nl.n_types,
                                          box_p, nl.r_cutoff, nl.perm.data_ptr(), nl.pos4.data_ptr(),
                                          nl.row_off.data_ptr(), nl.nbr.data_ptr(), nl.rev.data_ptr(), nl.cap, g,
                                          self.totals.data_ptr(), int(self.row_bound), nl.ws.data_ptr(),
                                          nl.ws.numel(), st),
                   "pt_nbr_build_async")
        arrays = (_lib.PermuteArray * 6)(*[_lib.PermuteArray(b.data_ptr(), w, wb) for b, w, wb in (
            (self.vel, 3, 8), (self.force, 3, 8), (self.mass, 1, 8), (self.inv_mass, 1, 8), (self.orig, 1, 4),
            (self.types, 1, 4))])
        _lib.check(lib.pt_permute_state_gated(self.n, nl.perm.data_ptr(), 6, arrays, self.tmp_state.data_ptr(),
                                              self.tmp_state.numel(), g, st), "pt_permute_state_gated")
        if self.skin > 0.0:
            _lib.check(lib.pt_copy_gated(4 * self.n, 8, nl.pos4.data_ptr(), self.ref_pos.data_ptr(), g, st),
                       "pt_copy_gated")
        # the descriptor kernels stage rows up to the bound; buffers are sized for the capacity
        nl.max_row = self.row_bound
        nl.n_pairs = nl.cap

    def lists_ok(self) -> bool:
        """Whether every gated build since the last check fitted the row bound and the
        list arrays (a host read: call where the host synchronises anyway)."""
        if not self.gated:
            return True
        t = self.totals.cpu().numpy()
        self._builds += int(t[3])
        ok = t[1] <= self.row_bound and t[2] <= self._nl.cap
        self.totals[1:].zero_()
        return bool(ok)

    def grow_lists(self) -> None:
        """Larger row bound and list arrays after an overflow (synchronous rebuild)."""
        self._rebuild_sync()
        self._size_bounds(grow=True)

    def _stale(self) -> bool:
        if self.skin <= 0.0:
            return True
        lib = _lib.load()
        box = host_box(self.box)
        _lib.check(lib.pt_max_displacement(self.n, self._nl.pos4.data_ptr(), self.ref_pos.data_ptr(),
                                           ctypes.cast(box, ctypes.c_void_p) if box is not None else None,
                                           self.disp.data_ptr(), _lib.stream_ptr()), "pt_max_displacement")
        d1, d2 = self.disp[:2].tolist()   # the two largest displacements
        return d1 + d2 > self.skin

    # ---- state snapshots (rewind) -------------------------------------------------
    def save_state(self) -> dict:
        return {"pos": self._nl.pos4.clone(), "vel": self.vel.clone(), "force": self.force.clone(),
                "orig": self.orig.clone(), "types": self.types.clone(), "mass": self.mass.clone(),
                "inv_mass": self.inv_mass.clone()}

    def restore_state(self, st: dict) -> None:
        self.orig = st["orig"].clone()
        self.types = st["types"].clone()
        self.vel = st["vel"].clone()
        self.force = st["force"].clone()
        self.mass = st["mass"].clone()
        self.inv_mass = st["inv_mass"].clone()
        self._rebuild_sync(st["pos"].clone())

    # ---- forces --------------------------------------------------------------
    def compute(self, out, force_mode: str, with_energy: bool = False, field: Optional[_ForceField] = None):
        """Forces of ``field`` (default: the driver) in internal order into ``out``."""
        import torch

        f = field or self.ff
        tag = "" if len(self.fields) == 1 else f"{f.label}/"
        mark = lambda ph, st: self._mark(tag + ph, st)   # noqa: E731
        nl = self.list_for(f)
        if nl is self._nl:
            return f.compute(nl, out, force_mode, with_energy, mark)
        tmp = torch.empty_like(out)
        f.compute(nl, tmp, force_mode, with_energy, mark)
        return scatter_rows(nl.perm, tmp, out)

    # ---- one step ------------------------------------------------------------
    def step(self, dt: float, integrator: str, force_mode: str, thermostat: Optional[BrendsenThermostat]):
        """One MD step.  With ``use_graph`` (gated steps only), the step's launch sequence is
        captured once into a HIP graph and replayed; it is captured again whenever a buffer
        or a parameter of the step changes."""
        args = (dt, integrator, force_mode, thermostat)
        if not (self.gated and self.use_graph) or self.timers is not None:
            self._graph = None
            self._step_body(*args, swap=True)
            return
        import torch

        if self._graph is not None and self._graph_key == self._key(args):
            self._graph.replay()
            return
        self._graph = None
        self._step_body(*args, swap=False)      # this step, run eagerly (sizes every buffer)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._step_body(*args, swap=False)  # recorded, not run
        self._graph, self._graph_key = g, self._key(args)

    def _key(self, args) -> tuple:
        """Everything a captured step depends on: its parameters and every buffer address."""
        ff = self.ff
        nl = self._nl
        bufs = (nl.pos4, nl.perm, nl.row_off, nl.nbr, nl.rev, nl.ws, self.vel, self.force, self.force_new,
                self.mass, self.inv_mass, self.orig, self.types, self.ref_pos, self.pos_in, self.tmp3, self.tmp1,
                self.tmpi, self.tmp_state, self.flag, self.totals, self.disp, self.reducer.ws, self.reducer.out, ff.G, ff.dEdG,
                ff.e_atom, ff.f_self, ff.edge)
        return (args[:3], id(args[3]), id(ff), id(ff.mlps), nl.cap, self.row_bound,
                tuple(0 if b is None else b.data_ptr() for b in bufs))

    def _step_body(self, dt: float, integrator: str, force_mode: str, thermostat: Optional[BrendsenThermostat],
                   swap: bool):
        lib = _lib.load()
        inv_m = self.inv_mass.data_ptr() if integrator == "physical" else None
        box = host_box(self.box)
        box_p = ctypes.cast(box, ctypes.c_void_p) if box is not None else None
        _lib.check(lib.pt_verlet_positions(self.n, self._nl.pos4.data_ptr(), self.vel.data_ptr(),
                                           self.force.data_ptr(), inv_m, dt, box_p, _lib.stream_ptr()),
                   "pt_verlet_positions")
        self._mark("nbr_build", True)
        if self.gated:
            self._rebuild_gated()
            inv_m = self.inv_mass.data_ptr() if integrator == "physical" else None
        elif self._stale():
            self._rebuild_sync()
            inv_m = self.inv_mass.data_ptr() if integrator == "physical" else None
        self._mark("nbr_build", False)
        self.compute(self.force_new, force_mode)
        _lib.check(lib.pt_verlet_velocities(self.n, self.vel.data_ptr(), self.force.data_ptr(),
                                            self.force_new.data_ptr(), inv_m, dt, _lib.stream_ptr()),
                   "pt_verlet_velocities")
        if swap:
            self.force, self.force_new = self.force_new, self.force
        else:   # a replayed step keeps its buffers
            self.force.copy_(self.force_new)
        if thermostat is not None:
            ke = self.reducer.kinetic(self.vel, self.mass)
            _lib.check(lib.pt_berendsen_rescale(self.n, self.vel.data_ptr(), ke.data_ptr(), dt,
                                                thermostat.time_constant, thermostat.target_temperature,
                                                KB, _lib.stream_ptr()), "pt_berendsen_rescale")

    def advance(self, steps: int, dt: float, integrator: str, force_mode: str,
                thermostat: Optional[BrendsenThermostat]) -> None:
        """``steps`` steps.  Gated steps are verified every ``check_every`` steps (and
        whenever the host reads the state): a stretch whose lists overflowed is redone
        from its checkpoint with larger bounds."""
        args = (dt, integrator, force_mode, thermostat)
        for _ in range(int(steps)):
            if self._args is not None and self._args != args:
                self.verify()
            if self.gated and self._since == 0:
                self._ck = self.save_state()
            self._args = args
            self.step(*args)
            if self.gated:
                self._since += 1
                if self._since >= self.check_every:
                    self.verify()

    def verify(self) -> None:
        """Check the gated builds since the last checkpoint (a host read) and redo the
        steps since then with larger bounds if a list overflowed."""
        if not self.gated or self._since == 0:
            return
        k = self._since
        self._since = 0
        while not self.lists_ok():
            self.redos += 1
            self.restore_state(self._ck)
            self.grow_lists()
            for _ in range(k):
                self.step(*self._args)

    # ---- views in input order ----------------------------------------------------
    def to_input(self, internal):
        import torch

        out = torch.empty_like(internal)
        return scatter_rows(self.orig, internal.reshape(internal.shape[0], -1), out.reshape(out.shape[0], -1))

    def positions(self):
        return self.to_input(self._nl.pos4[:, :3].contiguous())


class System:
    """Structure + potential + velocities + masses (system.py:55-164).

    Positions, velocities and forces live on the device in the MD engine;
    the properties return them in input atom order.
    """

    def __init__(self, potential: NeuralNetworkPotential, structure, velocities, masses,
                 force_mode: str = "reference", precision: str = "float64", skin: float = 0.0,
                 sync_free: bool = True, graph: bool = False):
        if force_mode not in ("reference", "full"):
            raise ValueError(f"unknown force mode '{force_mode}'")
        self.sync_free = bool(sync_free)
        self.graph = bool(graph)
        self.potential = potential
        self.structure = structure
        self._velocities = velocities
        self.masses = masses
        self.force_mode = force_mode
        self.precision = precision
        self.skin = float(skin)
        self._md: Optional[_DeviceMD] = None
        self.update_forces_from_positions()

    @property
    def velocities(self):
        if self._md is not None:
            self._md.verify()
            self._velocities = self._md.to_input(self._md.vel)
        return self._velocities

    @velocities.setter
    def velocities(self, value) -> None:
        import torch

        value = torch.as_tensor(value, dtype=torch.float64, device=self.masses.device).reshape(-1, 3).contiguous()
        self._velocities = value
        if self._md is not None:
            self._md.verify()
            self._md.vel = gather_rows(self._md.orig, value)

    @classmethod
    def from_structure(cls, structure, potential, temperature: float = 300.0, seed: int = 2024,
                       force_mode: str = "reference", precision: Optional[str] = None,
                       skin: float = 0.0, sync_free: bool = True, graph: bool = False) -> "System":
        """(system.py:75-96).  sync_free: periodic steps never wait for the host;
        graph: such steps are replayed from a captured HIP graph."""
        masses = structure.element_map.get_masses_from_structure(structure)
        vel = cls.generate_random_velocities(temperature, masses, seed)
        prec = precision or getattr(potential, "precision", "float64")
        return cls(potential, structure.copy(), vel, masses, force_mode, prec, skin, sync_free, graph)

    @staticmethod
    def generate_random_velocities(temperature: float, masses, seed: int):
        """Maxwell-Boltzmann velocities (system.py:83-96): normal, rescaled to T, COM removed."""
        import torch

        m = masses.detach().cpu().numpy().reshape(-1, 1)
        rng = np.random.default_rng(seed)
        v = rng.normal(size=(m.shape[0], 3))
        t_cur = 2.0 * (0.5 * np.sum(m * v * v)) / (3 * m.shape[0] * KB)
        v *= np.sqrt(temperature / t_cur)
        v -= np.sum(m * v, axis=0) / np.sum(m)
        return torch.as_tensor(v, dtype=torch.float64, device=masses.device)

    # ---- device state ------------------------------------------------------------
    def _engine(self, extra=()) -> _DeviceMD:
        if self._md is None:
            self._md = _DeviceMD(self, self.skin, extra, sync_free=self.sync_free)
            self._md.use_graph = self.graph
        return self._md

    def _sync_from_engine(self) -> None:
        md = self._md
        if md is None:
            return
        md.verify()
        self.structure.positions = md.positions().to(self.structure.dtype)
        self.structure.forces = md.to_input(md.force)

    def update_forces_from_positions(self) -> None:
        md = self._engine()
        md.compute(md.force, self.force_mode)
        self.structure.forces = md.to_input(md.force)

    def update_total_potential_energy_from_positions(self) -> None:
        self.structure.total_energy = self.get_potential_energy()

    # ---- observables (system.py:115-140) -----------------------------------------
    def get_potential_energy(self):
        md = self._engine()
        md.verify()
        md.compute(md.force_new, self.force_mode, with_energy=True)
        return md.energy.clone()

    def get_kinetic_energy(self):
        md = self._engine()
        md.verify()
        return md.reducer.kinetic(md.vel, md.mass).clone()

    def get_total_energy(self):
        return self.get_potential_energy() + self.get_kinetic_energy()

    def get_temperature(self):
        return 2.0 * self.get_kinetic_energy() / (3 * self.natoms * KB)

    def get_pressure(self):
        """(2 KE + sum x . F) / (3 V) (system.py:122-129, 205-211)."""
        if self.structure.box is None:
            raise AssertionError("Calculating pressure... input structure must have PBC box")
        self._sync_from_engine()
        vir = 2.0 * self.get_kinetic_energy() + (self.structure.positions * self.structure.forces).sum()
        return vir / (3.0 * self.structure.box.volume)

    def get_center_of_mass_velocity(self):
        self._sync_from_engine()
        m = self.masses
        return (m * self.velocities).sum(0) / m.sum()

    def get_center_of_mass_position(self):
        self._sync_from_engine()
        m = self.masses
        return (m * self.structure.positions).sum(0) / m.sum()

    @property
    def positions(self):
        self._sync_from_engine()
        return self.structure.positions

    @property
    def forces(self):
        self._sync_from_engine()
        return self.structure.forces

    @property
    def box(self):
        return self.structure.box

    @property
    def natoms(self) -> int:
        return self.structure.natoms


class MDSimulator:
    """Velocity-Verlet MD (molecular_dynamics.py:33-90)."""

    def __init__(self, time_step: float, thermostat: Optional[BrendsenThermostat] = None,
                 integrator: str = "reference"):
        if integrator not in ("reference", "physical"):
            raise ValueError(f"unknown integrator '{integrator}'")
        self.time_step = float(time_step)
        self.thermostat = thermostat
        self.integrator = integrator
        self.step = 0
        self.elapsed_time = 0.0

    def simulate_one_step(self, system: System) -> None:
        system._engine().advance(1, self.time_step, self.integrator, system.force_mode, self.thermostat)
        self.step += 1
        self.elapsed_time += self.time_step

    def run(self, system: System, num_steps: int) -> None:
        md = system._engine()
        md.advance(num_steps, self.time_step, self.integrator, system.force_mode, self.thermostat)
        self.step += num_steps
        self.elapsed_time += num_steps * self.time_step

    def repr_physical_params(self, system: System) -> str:
        if system.structure.box is None:
            return ""
        return (f"{self.step:<10} "
                f"time[ps]:{units.TO_PICO_SECOND * self.elapsed_time:<10.5f} "
                f"Temp[K]:{float(system.get_temperature()):<10.5f} "
                f"Etot[Ha]:{float(system.get_total_energy()):<15.10f} "
                f"Epot[Ha]:{float(system.get_potential_energy()):<15.10f} "
                f"Pres[kb]:{float(system.get_pressure()) * units.TO_KILO_BAR:<10.5f}")


def _write_xyz(path: Path, system: System, append: bool) -> None:
    pos = system.positions.detach().cpu().numpy() * units.TO_ANGSTROM
    els = system.structure.get_elements()
    with open(path, "a" if append else "w") as f:
        f.write(f"{len(els)}\n")
        if system.structure.box is not None:
            L = np.asarray(system.structure.box.length_host) * units.TO_ANGSTROM
            f.write(f'Lattice="{L[0]} 0 0 0 {L[1]} 0 0 0 {L[2]}" Properties=species:S:1:pos:R:3 pbc="T T T"\n')
        else:
            f.write("Properties=species:S:1:pos:R:3\n")
        for e, p in zip(els, pos):
            f.write(f"{e} {p[0]:.10f} {p[1]:.10f} {p[2]:.10f}\n")


def simulate(system: System, simulator: MDSimulator, num_steps: int = 1, output_freq: Optional[int] = None,
             filename=None, append: bool = False) -> None:
    """Run ``num_steps`` steps, printing every ``output_freq`` (simulate.py:31-88).

    Configurations are written as extended XYZ (the reference writes through
    ASE) when ``filename`` is given.
    """
    if output_freq is None:
        output_freq = 1 if num_steps < 100 else int(0.01 * num_steps)
    is_output = output_freq > 0
    path = Path(filename) if filename is not None else None
    if path is not None and not append:
        path.write_text("")
    init_step = simulator.step
    try:
        for _ in range(num_steps):
            if is_output and ((simulator.step - init_step) % output_freq == 0):
                print(simulator.repr_physical_params(system))
                if path is not None:
                    _write_xyz(path, system, append=True)
            simulator.simulate_one_step(system)
    except KeyboardInterrupt:
        print("KeyboardInterrupt")
    if is_output:
        print(simulator.repr_physical_params(system))
