"""High-dimensional NNP (potentials/nnp/potential.py:30-389) on gfx950.

Public surface (reference signatures):
  NeuralNetworkPotential.from_runner(path)      potential.py:47-65
  .load_scaler() / .load_model() / .load()      potential.py:104-137
  .__call__(structure) -> total energy          potential.py:67-81
  .compute_forces(structure) -> (N, 3)          potential.py:83-102
Extensions:
  precision = "float64" | "mixed"               descriptor/MLP arithmetic
  compute_forces(structure, mode="reference")   the reference's self term
                                                F_i = -dE_i/dr_i (force.py:16-43)
  compute_forces(structure, mode="full")        F = -grad E
  energy_and_forces(structure, mode)            both from one evaluation

One evaluation = neighbour build (pt_nbr_build) -> per element descriptors
(pt_acsf_forward) -> scaler + MLP + dE/dG (pt_mlp_energy_grad) -> FP64 energy
sum (pt_reduce_sum_f64) -> adjoint (pt_acsf_adjoint) -> scatter-free force
gather (pt_force_gather).
"""

from __future__ import annotations

import json
from collections import defaultdict
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .atoms import ElementMap
from .descriptors import (ACSF, G1, G2, G3, G9, CutoffFunction, DescriptorScaler, NeighborElements,
                          ScalerParams)
from .engine import (NeighborList, Reducer, acsf_adjoint, acsf_forward, force_gather, scatter_rows)
from .models import MLP_MODES, DeviceMLP, ModelParams, NeuralNetworkModel, UniformInitializer

# n2p2 / RuNNer keyword codes (settings.py:43-80)
_CUTOFF_BY_CODE = {"0": "hard", "1": "cos", "2": "tanhu", "3": "tanh", "4": "exp", "5": "poly1", "6": "poly2"}
_ACT_BY_CODE = {"l": "identity", "t": "tanh", "s": "logistic", "p": "softplus", "r": "relu",
                "g": "gaussian", "c": "cos", "e": "exp", "h": "harmonic"}
_SCALER_BY_KEY = {"center_symmetry_functions": "center", "scale_symmetry_functions": "scale",
                  "scale_center_symmetry_functions": "scale_center",
                  "scale_center_symmetry_functions_sigma": "scale_center_sigma"}
_UPDATER_BY_CODE = {"0": "gradient_descent", "1": "kalman_filter"}
_FLOAT_KEYS = ("force_weight", "short_force_fraction", "short_energy_fraction", "kalman_epsilon", "kalman_q0",
               "kalman_qtau", "kalman_qmin", "kalman_eta", "kalman_etatau", "kalman_etamax", "kalman_lambda_short",
               "kalman_neu_short")

PRECISION_MODES = {
    # name: (descriptor precision, MLP mode)
    "float64": (_lib.PT_F64, _lib.PT_MLP_F64),
    "mixed": (_lib.PT_F32, _lib.PT_MLP_BF16X3),
    "mixed_bf16": (_lib.PT_F32, _lib.PT_MLP_BF16),
}


@dataclass
class SymFuncArgs:
    central_element: str
    acsf_type: int
    neighbor_element_j: str
    eta: float
    r_cutoff: float
    r_shift: float = 0.0
    neighbor_element_k: Optional[str] = None
    lambda0: Optional[float] = None
    zeta: Optional[float] = None


@dataclass
class NNPSettings:
    """The keywords of potentials/nnp/settings.py:83-137 that the potential uses."""

    elements: List[str]
    global_nodes_short: List[int]
    global_activation_short: List[str]
    symfunction_short: List[SymFuncArgs]
    random_seed: int = 2023
    atom_energy: Dict[str, float] = field(default_factory=dict)
    scaler_save_format: str = "scaling.{:03d}.json"
    model_save_format: str = "weights.{:03d}.npz"
    weights_min: float = -1.0
    weights_max: float = 1.0
    cutoff_type: str = "tanh"
    scale_type: str = "center"
    scale_min_short: float = 0.0
    scale_max_short: float = 1.0
    # trainer keywords (settings.py:106-131)
    epochs: int = 1
    updater_type: str = "gradient_descent"
    force_weight: float = 1.0
    short_force_fraction: float = 0.1
    short_energy_fraction: float = 1.0
    kalman_type: int = 0
    kalman_epsilon: float = 0.01
    kalman_q0: float = 0.01
    kalman_qtau: float = 2.302
    kalman_qmin: float = 1.0e-6
    kalman_eta: float = 0.01
    kalman_etatau: float = 2.302
    kalman_etamax: float = 1.0
    kalman_lambda_short: float = 0.96
    kalman_neu_short: float = 0.9995
    extra: Dict[str, Any] = field(default_factory=dict)

    @classmethod
    def from_file(cls, filename) -> "NNPSettings":
        p = Path(filename)
        if p.suffix == ".json":
            return cls.from_json(p)
        if p.suffix == ".nn":
            return cls.from_nn(p)
        raise ValueError(f"Unknown file format '{p.suffix}': {p}")

    @classmethod
    def from_json(cls, filename) -> "NNPSettings":
        """JSON settings; symmetry functions in NamedTuple order (settings.py:15-37)."""
        d = json.loads(Path(filename).read_text())
        sfs = []
        for a in d.get("symfunction_short", []):
            if int(a[1]) < 3:
                c, t, j, eta, rc, rs = a
                sfs.append(SymFuncArgs(c, int(t), j, float(eta), float(rc), float(rs)))
            else:
                c, t, j, eta, rc, rs, k, lam, zeta = a
                sfs.append(SymFuncArgs(c, int(t), j, float(eta), float(rc), float(rs), k, float(lam), float(zeta)))
        known = {f for f in cls.__dataclass_fields__}
        kw = {k: v for k, v in d.items() if k in known and k != "symfunction_short"}
        # the reference's JSON fixture spells the save formats "*_naming_format"
        kw.setdefault("scaler_save_format", d.get("scaler_save_naming_format", "scaling.{:03d}.json"))
        kw["symfunction_short"] = sfs
        s = cls(**kw)
        s.extra = {k: v for k, v in d.items() if k not in known}
        return s

    @classmethod
    def from_nn(cls, filename) -> "NNPSettings":
        """RuNNer/n2p2 ``input.nn`` keywords (settings.py:154-308)."""
        kw: Dict[str, Any] = {"atom_energy": {}, "symfunction_short": []}
        for raw in Path(filename).read_text().splitlines():
            if raw.startswith("#"):
                continue
            line = raw.split("#", 1)[0]
            tok = line.split()
            if not tok:
                continue
            key, args = tok[0].lower(), tok[1:]
            if key == "elements":
                kw["elements"] = sorted(set(args), key=ElementMap.get_atomic_number_from_element)
            elif key == "atom_energy":
                kw["atom_energy"][args[0]] = float(args[1])
            elif key == "random_seed":
                kw["random_seed"] = int(args[0])
            elif key == "global_nodes_short":
                kw["global_nodes_short"] = [int(t) for t in args]
            elif key == "global_activation_short":
                kw["global_activation_short"] = [_ACT_BY_CODE[t] for t in args]
            elif key in ("weights_min", "weights_max", "scale_min_short", "scale_max_short") or key in _FLOAT_KEYS:
                kw[key] = float(args[0])
            elif key in ("epochs", "kalman_type"):
                kw[key] = int(args[0])
            elif key == "updater_type":
                kw[key] = _UPDATER_BY_CODE[args[0]]
            elif key == "cutoff_type":
                kw["cutoff_type"] = _CUTOFF_BY_CODE[args[0]]
            elif key in _SCALER_BY_KEY:
                kw["scale_type"] = _SCALER_BY_KEY[key]
            elif key == "symfunction_short":
                t = int(args[1])
                if t < 3:
                    kw["symfunction_short"].append(
                        SymFuncArgs(args[0], t, args[2], float(args[3]), float(args[5]), float(args[4])))
                else:
                    kw["symfunction_short"].append(SymFuncArgs(
                        args[0], t, args[2], float(args[4]), float(args[7]),
                        float(args[8]) if len(args) == 9 else 0.0, args[3], float(args[5]), float(args[6])))
        return cls(**kw)


@dataclass(frozen=True)
class AtomicPotential:
    """Descriptor -> scaler -> model of one element (atomic_potential.py:11-60)."""

    descriptor: ACSF
    scaler: DescriptorScaler
    model: NeuralNetworkModel

    @property
    def model_input_size(self) -> int:
        return self.descriptor.num_symmetry_functions


class _Evaluation:
    """Device buffers and launch sequence for one structure shape."""

    def __init__(self, pot: "NeuralNetworkPotential", structure, precision: str):
        import torch

        self.dev = _lib.require_device()
        self.precision = precision
        self.g_prec, self.mlp_mode = PRECISION_MODES[precision]
        self.emap = structure.element_map
        self.n = structure.natoms
        self.nl = NeighborList(pot.r_cutoff, self.emap.n_types)
        self.ld = max(ap.model_input_size for ap in pot.atomic_potentials.values())
        fdt = torch.float64 if self.g_prec == _lib.PT_F64 else torch.float32
        self.G = torch.zeros((self.n, self.ld), dtype=fdt, device=self.dev)
        self.dEdG = torch.zeros((self.n, self.ld), dtype=fdt, device=self.dev)
        self.e_atom = torch.zeros(self.n, dtype=torch.float64, device=self.dev)
        self.f_self = torch.zeros((self.n, 3), dtype=torch.float64, device=self.dev)
        self.f_full = torch.zeros((self.n, 3), dtype=torch.float64, device=self.dev)
        self.edge = None
        self.reducer = Reducer(self.n, self.dev)
        self.tables = {}
        self.covered = True
        for e in pot.elements:
            if e in self.emap.element_to_atom_type:
                self.tables[e] = pot.atomic_potentials[e].descriptor.device_table(self.emap, self.dev)
        for e in self.emap.unique_elements:
            if e not in pot.atomic_potentials:
                self.covered = False
        self.mlps = pot._device_mlps(self.dev)

    def matches(self, structure, precision: str) -> bool:
        return (structure.natoms == self.n and structure.element_map == self.emap and precision == self.precision)

    def build(self, structure) -> None:
        self.nl.build(structure.positions, structure.atom_types, structure.box_lengths,
                      structure.type_counts(), with_rev=True)

    def _blocks(self, pot):
        for e in pot.elements:
            t = self.emap.element_to_atom_type.get(e)
            if t is None:
                continue
            b, cnt = self.nl.type_block(t)
            if cnt > 0:
                yield e, b, cnt

    def energy(self, pot):
        for e, b, cnt in self._blocks(pot):
            acsf_forward(self.nl, self.tables[e], self.g_prec, cnt, center_begin=b, out=self.G[b:], ld=self.ld)
            self.mlps[e].energy_grad(self.mlp_mode, cnt, self.G[b:], self.ld, self.g_prec,
                                     self.e_atom[b:], self.dEdG[b:])
        return self.reducer.sum(self.e_atom)

    def forces(self, pot, full: bool):
        import torch

        need = self.nl.n_pairs * 3
        edt = torch.float64 if self.g_prec == _lib.PT_F64 else torch.float32
        if full and (self.edge is None or self.edge.numel() < need or self.edge.dtype != edt):
            self.edge = torch.zeros(max(need, 3), dtype=edt, device=self.dev)
        if full and not self.covered:
            self.edge.zero_()
        self.f_self.zero_()
        for e, b, cnt in self._blocks(pot):
            acsf_adjoint(self.nl, self.tables[e], self.g_prec, cnt, self.dEdG[b:], self.ld, center_begin=b,
                         f_self=self.f_self[b:], edge_grad=self.edge if full else None)
        if full:
            force_gather(self.nl, self.edge, self.g_prec, out=self.f_full)
            return self.f_full
        return self.f_self

    def to_input_order(self, internal):
        import torch

        out = torch.empty_like(internal)
        return scatter_rows(self.nl.perm, internal, out)


@dataclass
class NeuralNetworkPotential:
    directory: Path
    elements: Tuple[str, ...]
    atomic_potentials: Dict[str, AtomicPotential]
    models_params: Dict[str, ModelParams]
    scalers_params: Dict[str, Optional[ScalerParams]]
    scaler_save_format: str = "scaling.{:03d}.json"
    model_save_format: str = "weights.{:03d}.npz"
    precision: str = "float64"
    settings: Optional[NNPSettings] = None

    def __post_init__(self) -> None:
        self._eval: Optional[_Evaluation] = None
        self._mlp_cache: Dict[str, DeviceMLP] = {}
        self._mlp_key = None

    # ---- construction ----------------------------------------------------
    @classmethod
    def from_runner(cls, filename, precision: str = "float64") -> "NeuralNetworkPotential":
        settings = NNPSettings.from_file(filename)
        return cls.from_settings(settings, directory=Path(filename).parent, precision=precision)

    @classmethod
    def from_settings(cls, settings: NNPSettings, directory=".", precision: str = "float64"):
        radials, angulars = defaultdict(list), defaultdict(list)
        for a in settings.symfunction_short:
            cfn = CutoffFunction.from_type(settings.cutoff_type, a.r_cutoff)
            if a.acsf_type == 1:
                radials[a.central_element].append((G1(cfn), NeighborElements(a.neighbor_element_j)))
            elif a.acsf_type == 2:
                radials[a.central_element].append(
                    (G2(cfn, r_shift=a.r_shift, eta=a.eta), NeighborElements(a.neighbor_element_j)))
            elif a.acsf_type in (3, 9):
                G = G3 if a.acsf_type == 3 else G9
                # potential.py:226-232 passes r_shift=r_cutoff; angular terms never use r_shift
                angulars[a.central_element].append(
                    (G(cfn, eta=a.eta, zeta=a.zeta, lambda0=a.lambda0, r_shift=a.r_cutoff),
                     NeighborElements(a.neighbor_element_j, a.neighbor_element_k)))
        aps = {}
        for e in settings.elements:
            model = NeuralNetworkModel(
                hidden_layers=tuple(zip(settings.global_nodes_short, settings.global_activation_short[:-1])),
                output_layer=(1, settings.global_activation_short[-1]),
                kernel_initializer=UniformInitializer((settings.weights_min, settings.weights_max)),
            )
            aps[e] = AtomicPotential(
                ACSF(e, tuple(radials[e]), tuple(angulars[e])),
                DescriptorScaler.from_type(settings.scale_type, settings.scale_min_short, settings.scale_max_short),
                model,
            )
        params = {e: aps[e].model.init(settings.random_seed + i, aps[e].model_input_size)
                  for i, e in enumerate(settings.elements)}
        return cls(Path(directory), tuple(settings.elements), aps, params, {e: None for e in settings.elements},
                   settings.scaler_save_format, settings.model_save_format, precision, settings)

    # ---- persistence ---------------------------------------------------------
    def load_scaler(self) -> None:
        for e in self.elements:
            z = ElementMap.get_atomic_number_from_element(e)
            self.scalers_params[e] = DescriptorScaler.load(Path(self.directory, self.scaler_save_format.format(z)))
        self._mlp_key = None

    def load_model(self) -> None:
        for e in self.elements:
            z = ElementMap.get_atomic_number_from_element(e)
            fmt = self.model_save_format
            if fmt.endswith(".pkl"):
                fmt = fmt[:-4] + ".npz"
            self.models_params[e] = self.atomic_potentials[e].model.load(Path(self.directory, fmt.format(z)))
        self._mlp_key = None

    def load(self) -> None:
        self.load_scaler()
        self.load_model()

    def save(self, directory=None) -> None:
        d = Path(directory or self.directory)
        for e in self.elements:
            z = ElementMap.get_atomic_number_from_element(e)
            if self.scalers_params[e] is not None:
                DescriptorScaler.save(self.scalers_params[e], d / self.scaler_save_format.format(z))
            self.atomic_potentials[e].model.save(d / self.model_save_format.format(z), self.models_params[e])

    def fit_scaler(self, structures: Sequence) -> None:
        """Scaler statistics over the GPU descriptors of ``structures`` (trainer.py:68-88)."""
        for e in self.elements:
            params = None
            for s in structures:
                if e not in s.element_map.element_to_atom_type or s.select_host(e).size == 0:
                    continue
                x = self.atomic_potentials[e].descriptor(s, precision="float64")
                params = DescriptorScaler.fit(x) if params is None else DescriptorScaler.partial_fit(params, x)
            self.scalers_params[e] = params
        self._mlp_key = None

    def set_model_params(self, params: Dict[str, ModelParams]) -> None:
        self.models_params = dict(params)
        self._mlp_key = None

    # ---- evaluation --------------------------------------------------------
    def _check_scaler_params_exist(self) -> None:
        if any(v is None for v in self.scalers_params.values()):
            raise ValueError(
                f"Scaler parameters are not set yet for all the elements ({self.scalers_params})."
                "Try loading or fitting the scaler first.")

    def _device_mlps(self, device) -> Dict[str, DeviceMLP]:
        key = (id(self.models_params), str(device), tuple(id(v) for v in self.scalers_params.values()))
        if self._mlp_key != key:
            self._mlp_cache = {}
            for e in self.elements:
                ap = self.atomic_potentials[e]
                sm, add = ap.scaler.device_affine(self.scalers_params[e], device)
                self._mlp_cache[e] = DeviceMLP(ap.model, self.models_params[e], ap.model_input_size,
                                               sm.cpu().numpy(), add, device)
            self._mlp_key = key
        return self._mlp_cache

    def _evaluation(self, structure, precision: Optional[str] = None) -> _Evaluation:
        self._check_scaler_params_exist()
        prec = precision or self.precision
        if self._eval is None or not self._eval.matches(structure, prec):
            self._eval = _Evaluation(self, structure, prec)
        else:
            self._eval.mlps = self._device_mlps(self._eval.dev)
        self._eval.build(structure)
        return self._eval

    def __call__(self, structure, precision: Optional[str] = None):
        """Total energy (device FP64 scalar) (potential.py:67-81)."""
        ev = self._evaluation(structure, precision)
        return ev.energy(self).clone()

    def compute_forces(self, structure, mode: str = "reference", precision: Optional[str] = None):
        """Forces in input atom order (potential.py:83-102).

        mode="reference": the reference's self term -dE_i/dr_i (neighbours frozen);
        mode="full": -grad E.
        """
        return self.energy_and_forces(structure, mode, precision)[1]

    def energy_and_forces(self, structure, mode: str = "full", precision: Optional[str] = None):
        if mode not in ("reference", "full"):
            raise ValueError(f"unknown force mode '{mode}'")
        ev = self._evaluation(structure, precision)
        energy = ev.energy(self).clone()
        f = ev.forces(self, full=(mode == "full"))
        return energy, ev.to_input_order(f)

    def atomic_energies(self, structure, precision: Optional[str] = None):
        """Per-atom model outputs E_i in input order (atomic_potential.py:24-48)."""
        ev = self._evaluation(structure, precision)
        ev.energy(self)
        return ev.to_input_order(ev.e_atom.reshape(-1, 1)).reshape(-1)

    # ---- properties ----------------------------------------------------------
    @property
    def descriptors(self) -> Dict[str, ACSF]:
        return {e: ap.descriptor for e, ap in self.atomic_potentials.items()}

    @property
    def scalers(self) -> Dict[str, DescriptorScaler]:
        return {e: ap.scaler for e, ap in self.atomic_potentials.items()}

    @property
    def models(self) -> Dict[str, NeuralNetworkModel]:
        return {e: ap.model for e, ap in self.atomic_potentials.items()}

    @property
    def r_cutoff(self) -> float:
        return max(ap.descriptor.r_cutoff for ap in self.atomic_potentials.values())

    @property
    def num_elements(self) -> int:
        return len(self.elements)


NNP = NeuralNetworkPotential


class _EnergyFunction:
    """torch.autograd.Function: E(positions) with dE/dpositions = -F (the
    hand-written adjoint kernels, scatter-free gather)."""

    _fn = None

    @classmethod
    def get(cls):
        if cls._fn is None:
            import torch

            class NNPEnergy(torch.autograd.Function):
                @staticmethod
                def forward(ctx, positions, pot, structure, precision):
                    s = structure.copy()
                    s.positions = positions.detach().to(torch.float64)
                    e, f = pot.energy_and_forces(s, mode="full", precision=precision)
                    ctx.save_for_backward(f)
                    return e

                @staticmethod
                def backward(ctx, grad_out):
                    (f,) = ctx.saved_tensors
                    return -f * grad_out, None, None, None

            cls._fn = NNPEnergy
        return cls._fn


def energy(pot: "NeuralNetworkPotential", structure, positions=None, precision: Optional[str] = None):
    """Differentiable total energy: ``torch.autograd.grad(energy(pot, s, x), x) == -F``."""
    x = structure.positions if positions is None else positions
    return _EnergyFunction.get().apply(x, pot, structure, precision)
