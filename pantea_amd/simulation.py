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
skin (checked on the device every step: no pair can have moved by more);
the descriptor kernels re-apply the exact mask
(r <= rc, r > 0), so every step sees exactly the reference's neighbour set.
"""

from __future__ import annotations

import ctypes
from pathlib import Path
from typing import Optional

import numpy as np

from . import _lib
from .atoms import ElementMap, host_box
from .engine import NeighborList, Reducer, acsf_adjoint, acsf_forward, force_gather, gather_rows, scatter_rows
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
            acsf_adjoint(nl, tab, self.g_prec, cnt, self.dEdG[b:], self.ld, center_begin=b,
                         f_self=target[b:], edge_grad=self.edge if full else None)
        mark("acsf_adjoint", False)
        if full:
            mark("force_gather", True)
            force_gather(nl, self.edge, self.g_prec, out=out)
            mark("force_gather", False)
        return out


class _DeviceMD:
    """Device-resident MD state of one System (internal atom order).

    ``extra`` potentials (e.g. a surrogate and its judge) share the state and
    the neighbour list, whose cutoff is the largest of all plus the skin.
    """

    def __init__(self, system: "System", skin: float, extra=()):
        import torch

        pot = system.potential
        s = system.structure
        self.dev = s.device
        self.n = s.natoms
        self.emap = s.element_map
        self.box = s.box_lengths
        self.skin = float(skin)
        n = self.n
        self.fields = {(id(pot), system.precision): _ForceField(pot, system.precision, self.emap, n, self.dev)}
        for p, prec in extra:
            self.fields.setdefault((id(p), prec), _ForceField(p, prec, self.emap, n, self.dev))
        self.ff = self.fields[(id(pot), system.precision)]
        self.rc = max(f.rc for f in self.fields.values())
        self.nls = [NeighborList(self.rc + self.skin, self.emap.n_types),
                    NeighborList(self.rc + self.skin, self.emap.n_types)]
        self.cur = 0
        self.type_counts = s.type_counts()
        # state in input order -> internal order of the first build
        nl = self.nls[0].build(s.positions, s.atom_types, self.box, self.type_counts, with_rev=True)
        perm = nl.perm
        self.orig = perm.clone()                                     # internal -> input index
        self.types = gather_rows(perm, s.atom_types.to(torch.int32).reshape(-1, 1)).reshape(-1)
        self.vel = gather_rows(perm, system.velocities.contiguous())
        mass = system.masses.reshape(-1, 1).contiguous()
        self.mass = gather_rows(perm, mass).reshape(-1).contiguous()
        self.inv_mass = (1.0 / self.mass).contiguous()
        self.ref_pos = nl.pos4.clone()
        self.force = torch.zeros((n, 3), dtype=torch.float64, device=self.dev)
        self.force_new = torch.zeros_like(self.force)
        self.disp = torch.zeros(514, dtype=torch.float64, device=self.dev)
        self.reducer = Reducer(n, self.dev)
        self.n_builds = 1
        self.timers = None  # {phase: [(start_event, end_event), ...]} when profiling

    def add_field(self, pot, precision: str) -> _ForceField:
        """Evaluate ``pot`` at ``precision`` on this state too (grows the list cutoff if needed)."""
        key = (id(pot), precision)
        if key not in self.fields:
            self.fields[key] = _ForceField(pot, precision, self.emap, self.n, self.dev)
            rc = max(f.rc for f in self.fields.values())
            if rc > self.rc:
                self.rc = rc
                pos = self.nl.pos4.clone()
                self.nls = [NeighborList(rc + self.skin, self.emap.n_types),
                            NeighborList(rc + self.skin, self.emap.n_types)]
                self.cur = 0
                self._rebuild(pos)
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
        """Make ``field`` the driving potential of step()."""
        self.ff = field

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
        return self.nls[self.cur]

    # ---- neighbour maintenance ---------------------------------------------
    def _rebuild(self, positions=None) -> None:
        """New list (and internal order) from ``positions`` (current order; default: current positions)."""
        old = self.nls[self.cur]
        new = self.nls[1 - self.cur]
        new.build(old.pos4 if positions is None else positions, self.types, self.box, self.type_counts,
                  with_rev=True)
        p = new.perm
        self.orig = gather_rows(p, self.orig.reshape(-1, 1)).reshape(-1).contiguous()
        self.types = gather_rows(p, self.types.reshape(-1, 1)).reshape(-1).contiguous()
        self.vel = gather_rows(p, self.vel)
        self.force = gather_rows(p, self.force)
        self.mass = gather_rows(p, self.mass.reshape(-1, 1)).reshape(-1).contiguous()
        self.inv_mass = gather_rows(p, self.inv_mass.reshape(-1, 1)).reshape(-1).contiguous()
        self.ref_pos.copy_(new.pos4)
        self.cur = 1 - self.cur
        self.n_builds += 1

    def _stale(self) -> bool:
        if self.skin <= 0.0:
            return True
        lib = _lib.load()
        box = host_box(self.box)
        _lib.check(lib.pt_max_displacement(self.n, self.nl.pos4.data_ptr(), self.ref_pos.data_ptr(),
                                           ctypes.cast(box, ctypes.c_void_p) if box is not None else None,
                                           self.disp.data_ptr(), _lib.stream_ptr()), "pt_max_displacement")
        d1, d2 = self.disp[:2].tolist()   # the two largest displacements
        return d1 + d2 > self.skin

    # ---- state snapshots (rewind) -------------------------------------------------
    def save_state(self) -> dict:
        return {"pos": self.nl.pos4.clone(), "vel": self.vel.clone(), "force": self.force.clone(),
                "orig": self.orig.clone(), "types": self.types.clone(), "mass": self.mass.clone(),
                "inv_mass": self.inv_mass.clone()}

    def restore_state(self, st: dict) -> None:
        self.orig = st["orig"].clone()
        self.types = st["types"].clone()
        self.vel = st["vel"].clone()
        self.force = st["force"].clone()
        self.mass = st["mass"].clone()
        self.inv_mass = st["inv_mass"].clone()
        self._rebuild(st["pos"])

    # ---- forces --------------------------------------------------------------
    def compute(self, out, force_mode: str, with_energy: bool = False, field: Optional[_ForceField] = None):
        f = field or self.ff
        tag = "" if len(self.fields) == 1 else f"{f.label}/"
        return f.compute(self.nl, out, force_mode, with_energy, lambda ph, st: self._mark(tag + ph, st))

    # ---- one step ------------------------------------------------------------
    def step(self, dt: float, integrator: str, force_mode: str, thermostat: Optional[BrendsenThermostat]):
        lib = _lib.load()
        inv_m = self.inv_mass.data_ptr() if integrator == "physical" else None
        box = host_box(self.box)
        box_p = ctypes.cast(box, ctypes.c_void_p) if box is not None else None
        _lib.check(lib.pt_verlet_positions(self.n, self.nl.pos4.data_ptr(), self.vel.data_ptr(),
                                           self.force.data_ptr(), inv_m, dt, box_p, _lib.stream_ptr()),
                   "pt_verlet_positions")
        if self._stale():
            self._mark("nbr_build", True)
            self._rebuild()
            self._mark("nbr_build", False)
        self.compute(self.force_new, force_mode)
        _lib.check(lib.pt_verlet_velocities(self.n, self.vel.data_ptr(), self.force.data_ptr(),
                                            self.force_new.data_ptr(), inv_m, dt, _lib.stream_ptr()),
                   "pt_verlet_velocities")
        self.force, self.force_new = self.force_new, self.force
        if thermostat is not None:
            ke = self.reducer.kinetic(self.vel, self.mass)
            _lib.check(lib.pt_berendsen_rescale(self.n, self.vel.data_ptr(), ke.data_ptr(), dt,
                                                thermostat.time_constant, thermostat.target_temperature,
                                                KB, _lib.stream_ptr()), "pt_berendsen_rescale")

    # ---- views in input order ----------------------------------------------------
    def to_input(self, internal):
        import torch

        out = torch.empty_like(internal)
        return scatter_rows(self.orig, internal.reshape(internal.shape[0], -1), out.reshape(out.shape[0], -1))

    def positions(self):
        return self.to_input(self.nl.pos4[:, :3].contiguous())


class System:
    """Structure + potential + velocities + masses (system.py:55-164).

    Positions, velocities and forces live on the device in the MD engine;
    the properties return them in input atom order.
    """

    def __init__(self, potential: NeuralNetworkPotential, structure, velocities, masses,
                 force_mode: str = "reference", precision: str = "float64", skin: float = 0.0):
        if force_mode not in ("reference", "full"):
            raise ValueError(f"unknown force mode '{force_mode}'")
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
            self._velocities = self._md.to_input(self._md.vel)
        return self._velocities

    @velocities.setter
    def velocities(self, value) -> None:
        import torch

        value = torch.as_tensor(value, dtype=torch.float64, device=self.masses.device).reshape(-1, 3).contiguous()
        self._velocities = value
        if self._md is not None:
            self._md.vel = gather_rows(self._md.orig, value)

    @classmethod
    def from_structure(cls, structure, potential, temperature: float = 300.0, seed: int = 2024,
                       force_mode: str = "reference", precision: Optional[str] = None,
                       skin: float = 0.0) -> "System":
        masses = structure.element_map.get_masses_from_structure(structure)
        vel = cls.generate_random_velocities(temperature, masses, seed)
        prec = precision or getattr(potential, "precision", "float64")
        return cls(potential, structure.copy(), vel, masses, force_mode, prec, skin)

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
            self._md = _DeviceMD(self, self.skin, extra)
        return self._md

    def _sync_from_engine(self) -> None:
        md = self._md
        if md is None:
            return
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
        md.compute(md.force_new, self.force_mode, with_energy=True)
        return md.energy.clone()

    def get_kinetic_energy(self):
        md = self._engine()
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
        system._engine().step(self.time_step, self.integrator, system.force_mode, self.thermostat)
        self.step += 1
        self.elapsed_time += self.time_step

    def run(self, system: System, num_steps: int) -> None:
        md = system._engine()
        for _ in range(num_steps):
            md.step(self.time_step, self.integrator, system.force_mode, self.thermostat)
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
