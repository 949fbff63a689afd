"""ACSF descriptors and the descriptor scaler.

Host-side mirror of the reference's descriptor interface:
  CutoffFunction.from_type(type, r_cutoff)   descriptors/acsf/cutoff.py:44-60
  G1(cfn), G2(cfn, r_shift, eta)             descriptors/acsf/radial.py:24-61
  G3/G9(cfn, eta, zeta, lambda0, r_shift)    descriptors/acsf/angular.py:30-107
  NeighborElements(j, k=None)                descriptors/acsf/symmetry.py:28-35
  ACSF(central_element, radial, angular)     descriptors/acsf/acsf.py:26-160
    __call__(structure, atom_index=None)     acsf.py:47-86
    grad(structure, atom_index=None)         acsf.py:88-120
  DescriptorScaler                            descriptors/scaler.py:43-203
The numbers come from the gfx950 kernels (pt_acsf_forward for values,
pt_acsf_tangent for the self-Jacobian); the symmetry functions here only
carry parameters.
"""

from __future__ import annotations

import json
from dataclasses import dataclass
from pathlib import Path
from typing import Dict, NamedTuple, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .engine import PRECISIONS, DeviceTable, NeighborList, SFRow, acsf_forward, acsf_tangent, gather_rows


# ---------------------------------------------------------------------------
# symmetry functions (parameter carriers)
# ---------------------------------------------------------------------------

@dataclass(frozen=True)
class CutoffFunction:
    r_cutoff: float
    cutoff_type: str = "tanh"

    @classmethod
    def from_type(cls, cutoff_type: str, r_cutoff: float) -> "CutoffFunction":
        if cutoff_type not in _lib.CUTOFF_CODES:
            raise KeyError(cutoff_type)
        return cls(float(r_cutoff), cutoff_type)

    @property
    def code(self) -> int:
        return _lib.CUTOFF_CODES[self.cutoff_type]

    def __repr__(self) -> str:
        return f"CutoffFunction(r_cutoff={self.r_cutoff})"


class NeighborElements(NamedTuple):
    neighbor_j: str
    neighbor_k: Optional[str] = None


@dataclass(frozen=True)
class G1:
    cfn: CutoffFunction
    kind = 1

    @property
    def r_cutoff(self) -> float:
        return self.cfn.r_cutoff


@dataclass(frozen=True)
class G2:
    cfn: CutoffFunction
    r_shift: float
    eta: float
    kind = 2

    @property
    def r_cutoff(self) -> float:
        return self.cfn.r_cutoff


@dataclass(frozen=True)
class G3:
    cfn: CutoffFunction
    eta: float
    zeta: float
    lambda0: float
    r_shift: float = 0.0
    kind = 3

    @property
    def r_cutoff(self) -> float:
        return self.cfn.r_cutoff


@dataclass(frozen=True)
class G9(G3):
    kind = 9


RadialSF = Tuple[object, NeighborElements]
AngularSF = Tuple[object, NeighborElements]


# ---------------------------------------------------------------------------
# ACSF
# ---------------------------------------------------------------------------

@dataclass(frozen=True)
class ACSF:
    """Atom-centred symmetry functions of one central element.

    Column layout: radial functions in insertion order, then angular ones
    (acsf.py:163-202).
    """

    central_element: str
    radial_symmetry_functions: Tuple[RadialSF, ...] = ()
    angular_symmetry_functions: Tuple[AngularSF, ...] = ()

    def __post_init__(self) -> None:
        object.__setattr__(self, "radial_symmetry_functions", tuple(self.radial_symmetry_functions))
        object.__setattr__(self, "angular_symmetry_functions", tuple(self.angular_symmetry_functions))
        object.__setattr__(self, "_tables", {})

    @property
    def num_radial_symmetry_functions(self) -> int:
        return len(self.radial_symmetry_functions)

    @property
    def num_angular_symmetry_functions(self) -> int:
        return len(self.angular_symmetry_functions)

    @property
    def num_symmetry_functions(self) -> int:
        return self.num_radial_symmetry_functions + self.num_angular_symmetry_functions

    @property
    def r_cutoff(self) -> float:
        return max(sf.r_cutoff for sf, _ in self.radial_symmetry_functions + self.angular_symmetry_functions)

    # ---- tables ------------------------------------------------------------
    def rows(self, element_to_atom_type: Dict[str, int]) -> Tuple[SFRow, ...]:
        """Resolve neighbour elements to atom types (0 when absent from the structure)."""
        out = []
        col = 0
        for sf, ne in self.radial_symmetry_functions:
            ej = element_to_atom_type.get(ne.neighbor_j, 0)
            eta = getattr(sf, "eta", 0.0)
            rs = getattr(sf, "r_shift", 0.0)
            out.append(SFRow(sf.kind, ej, 0, float(eta), float(rs), 0.0, 0.0, sf.r_cutoff, sf.cfn.code, col))
            col += 1
        for sf, ne in self.angular_symmetry_functions:
            ej = element_to_atom_type.get(ne.neighbor_j, 0)
            ek = element_to_atom_type.get(ne.neighbor_k, 0)
            out.append(SFRow(sf.kind, ej, ek, float(sf.eta), float(sf.r_shift), float(sf.zeta),
                             float(sf.lambda0), sf.r_cutoff, sf.cfn.code, col))
            col += 1
        return tuple(out)

    def device_table(self, element_map, device) -> DeviceTable:
        key = (tuple(sorted(element_map.element_to_atom_type.items())), str(device))
        tab = self._tables.get(key)
        if tab is None:
            rows = [r for r in self.rows(element_map.element_to_atom_type)]
            # functions whose neighbour element is absent contribute exactly 0:
            # keep their columns (zero-filled) but drop them from the kernel table
            live = [r for r in rows if r.ej > 0 and (r.kind in (1, 2) or r.ek > 0)]
            tab = DeviceTable(live, device)
            tab.n_sf = self.num_symmetry_functions
            self._tables[key] = tab
        return tab

    # ---- evaluation ----------------------------------------------------------
    def _centers(self, structure, atom_index, check_element: bool) -> np.ndarray:
        if atom_index is None:
            return structure.select_host(self.central_element) if check_element else np.arange(structure.natoms)
        idx = np.atleast_1d(np.asarray(atom_index.cpu() if hasattr(atom_index, "cpu") else atom_index)).astype(np.int64)
        if check_element:
            t = structure.element_map.element_to_atom_type.get(self.central_element, -1)
            if not np.all(structure.atom_types_host[idx] == t):
                raise ValueError(
                    f"Inconsistent central element '{self.central_element}': input atom index={atom_index}")
        elif not np.all((idx >= 0) & (idx < structure.natoms)):
            raise ValueError(f"unexpected atom_index={atom_index}. Input index must be between [0, {structure.natoms})")
        return idx

    def _prepare(self, structure, precision: str):
        import torch

        dev = _lib.require_device()
        nl = NeighborList(self.r_cutoff, structure.element_map.n_types)
        nl.build(structure.positions, structure.atom_types, structure.box_lengths,
                 structure.type_counts(), with_rev=False)
        tab = self.device_table(structure.element_map, dev)
        inv = torch.as_tensor(nl.inverse_perm_host().astype(np.int32), device=dev)
        return nl, tab, inv, PRECISIONS[precision]

    def __call__(self, structure, atom_index=None, precision: str = "float64"):
        """Descriptor values (n, n_sf) of the selected central atoms (acsf.py:47-86)."""
        import torch

        if self.num_symmetry_functions == 0:
            import warnings

            warnings.warn("No symmetry function was found")
        idx = self._centers(structure, atom_index, check_element=True)
        nl, tab, inv, prec = self._prepare(structure, precision)
        centers = inv[torch.as_tensor(idx, device=inv.device)].contiguous()
        return acsf_forward(nl, tab, prec, int(centers.numel()), centers=centers)

    def grad(self, structure, atom_index=None, precision: str = "float64"):
        """Self-Jacobian dG_i/dr_i, shape (n, n_sf, 3) (acsf.py:88-120, 215-228)."""
        import torch

        idx = self._centers(structure, atom_index, check_element=False)
        nl, tab, inv, prec = self._prepare(structure, precision)
        centers = inv[torch.as_tensor(idx, device=inv.device)].contiguous()
        n = int(centers.numel())
        out = torch.empty((n, self.num_symmetry_functions, 3),
                          dtype=torch.float64 if prec == _lib.PT_F64 else torch.float32, device=inv.device)
        for d in range(3):
            delta = torch.zeros((structure.natoms, 3), dtype=torch.float64, device=inv.device)
            delta[:, d] = 1.0
            out[:, :, d] = acsf_tangent(nl, tab, prec, n, delta, center_only=True, centers=centers)
        return out

    def __repr__(self) -> str:
        return f"ACSF(central_element='{self.central_element}', num_symmetry_functions={self.num_symmetry_functions})"


# ---------------------------------------------------------------------------
# scaler (descriptors/scaler.py)
# ---------------------------------------------------------------------------

class ScalerParams(NamedTuple):
    dimension: int
    nsamples: int
    mean: np.ndarray
    sigma: np.ndarray
    minval: np.ndarray
    maxval: np.ndarray


class ScalerWarnings(NamedTuple):
    number_of_warnings: int
    max_number_of_warnings: int


_SCALE_TYPES = ("center", "scale", "scale_center", "scale_center_sigma")


def _on_device(data) -> bool:
    return hasattr(data, "is_cuda") and bool(data.is_cuda)


def _device_rows(data):
    """(tensor, rows, cols, ld, precision) of a 1-D/2-D FP64/FP32 device tensor."""
    import torch

    x = data if data.dim() == 2 else data.reshape(1, -1)
    if x.dtype not in (torch.float64, torch.float32):
        x = x.to(torch.float64)
    if x.stride(1) != 1:
        x = x.contiguous()
    prec = _lib.PT_F64 if x.dtype == torch.float64 else _lib.PT_F32
    return x, int(x.shape[0]), int(x.shape[1]), int(x.stride(0)), prec


def params_to_device(params: "ScalerParams", device):
    """(4, d) FP64 device tensor mean | sigma | minval | maxval (the layout of pt_scaler_*)."""
    import torch

    return torch.as_tensor(np.stack([np.asarray(params.mean, np.float64), np.asarray(params.sigma, np.float64),
                                     np.asarray(params.minval, np.float64), np.asarray(params.maxval, np.float64)]),
                           dtype=torch.float64, device=device).contiguous()


def _device_partial_fit(data, params: Optional["ScalerParams"]) -> "ScalerParams":
    """fit (params None) or partial_fit on device rows with pt_scaler_partial_fit."""
    import torch

    x, n, d, ld, prec = _device_rows(data)
    lib = _lib.load()
    if params is None:
        dev_p = torch.zeros((4, d), dtype=torch.float64, device=x.device)
        m = 0
    else:
        if int(params.dimension) != d:
            raise ValueError(f"scaler dimension {params.dimension} != data width {d}")
        dev_p = params_to_device(params, x.device)
        m = int(params.nsamples)
    ws = torch.empty(lib.pt_scaler_workspace_bytes(n, d), dtype=torch.uint8, device=x.device)
    _lib.check(lib.pt_scaler_partial_fit(n, d, x.data_ptr(), ld, prec, float(m), dev_p.data_ptr(), ws.data_ptr(),
                                         ws.numel(), _lib.stream_ptr()), "pt_scaler_partial_fit")
    h = dev_p.cpu().numpy()
    return ScalerParams(d, m + n, h[0].copy(), h[1].copy(), h[2].copy(), h[3].copy())


def extrapolation_counts(data, params_dev, counters, margin: float = 0.0) -> None:
    """Accumulate the out-of-range rows/values of device rows ``data`` into the
    device counters (3 x uint64, pt_scaler_extrapolation); no synchronisation."""
    x, n, d, ld, prec = _device_rows(data)
    _lib.check(_lib.load().pt_scaler_extrapolation(n, d, x.data_ptr(), ld, prec, params_dev.data_ptr(),
                                                   float(margin), counters.data_ptr(), _lib.stream_ptr()),
               "pt_scaler_extrapolation")


class DescriptorScaler:
    """Affine descriptor scaling (scaler.py:43-203, transforms :206-246).

    Every scaler kind is the affine map x' = (x - sub) * mul + add; the MLP
    kernel applies it on the fly, so only (sub, mul, add) reach the device.
    Statistics (fit / partial_fit, scaler.py:249-283) and the extrapolation
    check (scaler.py:112-142, 286-295) of device descriptor rows run on the
    GPU (pt_scaler_partial_fit, pt_scaler_extrapolation); host arrays are
    reduced with numpy in FP64.  The parameters themselves are small host
    arrays.
    """

    def __init__(self, scale_type: str = "center", scale_min: float = 0.0, scale_max: float = 1.0):
        if scale_type not in _SCALE_TYPES:
            raise KeyError(scale_type)
        if not scale_min < scale_max:
            raise ValueError("Unexpected scale range values")
        self.scale_type = scale_type
        self.scale_min = float(scale_min)
        self.scale_max = float(scale_max)

    @classmethod
    def from_type(cls, scale_type: str, scale_min: float = 0.0, scale_max: float = 1.0) -> "DescriptorScaler":
        return cls(scale_type, scale_min, scale_max)

    @staticmethod
    def _np(data) -> np.ndarray:
        if hasattr(data, "detach"):
            data = data.detach().cpu().numpy()
        return np.atleast_2d(np.asarray(data, dtype=np.float64))

    @classmethod
    def fit(cls, data) -> ScalerParams:
        if _on_device(data):
            return _device_partial_fit(data, None)
        x = cls._np(data)
        return ScalerParams(x.shape[1], x.shape[0], x.mean(0), x.std(0), x.min(0), x.max(0))

    @classmethod
    def partial_fit(cls, params: ScalerParams, data) -> ScalerParams:
        if _on_device(data):
            return _device_partial_fit(data, params)
        x = cls._np(data)
        m, n = params.nsamples, x.shape[0]
        new_mean, new_sigma = x.mean(0), x.std(0)
        fm, fn = m / (m + n), n / (m + n)
        diff = params.mean - new_mean
        sigma = np.sqrt((fm * params.sigma) * params.sigma + (fn * new_sigma) * new_sigma
                        + (fm * diff) * (fn * diff))
        return ScalerParams(params.dimension, m + n, fm * params.mean + fn * new_mean, sigma,
                            np.minimum(params.minval, x.min(0)), np.maximum(params.maxval, x.max(0)))

    def affine(self, params: ScalerParams) -> Tuple[np.ndarray, np.ndarray, float]:
        """(sub, mul, add) with x' = (x - sub) * mul + add."""
        lo, hi = self.scale_min, self.scale_max
        mean = np.asarray(params.mean, dtype=np.float64)
        if self.scale_type == "center":
            return mean, np.ones_like(mean), 0.0
        rng = np.asarray(params.maxval, dtype=np.float64) - np.asarray(params.minval, dtype=np.float64)
        if self.scale_type == "scale":
            return np.asarray(params.minval, dtype=np.float64), (hi - lo) / rng, lo
        if self.scale_type == "scale_center":
            return mean, (hi - lo) / rng, lo
        return mean, (lo - hi) / np.asarray(params.sigma, dtype=np.float64), lo

    def device_affine(self, params: ScalerParams, device):
        import torch

        sub, mul, add = self.affine(params)
        return torch.as_tensor(np.stack([sub, mul]), dtype=torch.float64, device=device).contiguous(), add

    @staticmethod
    def initialize_warnings(number_of_warnings: int = 0, max_number_of_warnings: int = -1) -> ScalerWarnings:
        return ScalerWarnings(number_of_warnings, max_number_of_warnings)

    @classmethod
    def check_warnings(cls, params: ScalerParams, data, warnings: ScalerWarnings) -> ScalerWarnings:
        """Counts one warning per call with any value outside [min, max] (scaler.py:112-142,286-295)."""
        if warnings.max_number_of_warnings < 0:
            return warnings
        if _on_device(data):
            import torch

            counters = torch.zeros(3, dtype=torch.int64, device=data.device)
            extrapolation_counts(data, params_to_device(params, data.device), counters)
            out = int(counters[0].item()) > 0
        else:
            x = cls._np(data)
            out = bool(np.any((x > params.maxval[None, :]) | (x < params.minval[None, :])))
        new = ScalerWarnings(warnings.number_of_warnings + int(out), warnings.max_number_of_warnings)
        if new.number_of_warnings >= new.max_number_of_warnings:
            import warnings as _w

            _w.warn(f"Exceeding maximum number scaler extrapolation warnings: {new.number_of_warnings}"
                    f" (max={new.max_number_of_warnings})")
        return new

    @staticmethod
    def save(params: ScalerParams, filename) -> None:
        d = {k: (np.asarray(v).tolist()) for k, v in params._asdict().items()}
        Path(filename).write_text(json.dumps(d, indent=4))

    @staticmethod
    def load(filename) -> ScalerParams:
        d = json.loads(Path(filename).read_text())
        return ScalerParams(int(d["dimension"]), int(d["nsamples"]),
                            *(np.asarray(d[k], dtype=np.float64) for k in ("mean", "sigma", "minval", "maxval")))

    def __repr__(self) -> str:
        return f"DescriptorScaler(transform='{self.scale_type}', scale_range=({self.scale_min}, {self.scale_max}))"
