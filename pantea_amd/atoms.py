"""Atoms on the GPU: ElementMap, Box and Structure.

Host-side mirror of the reference's atoms layer (atoms/element.py:73-150,
atoms/box.py:14-109, atoms/structure.py:21-387) with torch tensors on the
HIP device as storage.  Positions are FP64 device tensors; atom types are
1-based int32, ordered by atomic number (element.py:88-117).  Wrapping into
the box on construction (structure.py:80-81) runs the pt_wrap_positions
kernel.
"""

from __future__ import annotations

import ctypes
from collections import defaultdict
from dataclasses import dataclass, field
from typing import Any, Dict, Iterator, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .units import units

# Symbols by atomic number.
_ELEMENTS: Tuple[str, ...] = (
    "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne",
    "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar", "K", "Ca",
    "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn",
    "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr", "Y", "Zr",
    "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn",
    "Sb", "Te", "I", "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd",
    "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb",
    "Lu", "Hf", "Ta", "W", "Re", "Os", "Ir", "Pt", "Au", "Hg",
    "Tl", "Pb", "Bi", "Po", "At", "Rn",
)
_ATOMIC_NUMBER: Dict[str, int] = {e: z for z, e in enumerate(_ELEMENTS, start=1)}

# Standard atomic weights (u), abridged to the precision the reference uses
# (element.py:28-58); masses in a.u. = weight / TO_ATOMIC_MASS.
_WEIGHT_U: Dict[str, float] = {
    "H": 1.008, "He": 4.003, "Li": 6.941, "Be": 9.012, "B": 10.811, "C": 12.011,
    "N": 14.007, "O": 15.999, "F": 18.998, "Ne": 20.180, "Na": 22.990, "Mg": 24.305,
    "Al": 26.982, "Si": 28.086, "P": 30.974, "S": 32.066, "Cl": 35.453, "Ar": 39.948,
    "K": 39.098, "Ca": 40.078, "Sc": 44.956, "Ti": 47.867, "V": 50.942, "Cr": 51.996,
    "Mn": 54.938, "Fe": 55.845, "Co": 58.933, "Ni": 58.693, "Cu": 63.546, "Zn": 65.38,
    "Ga": 69.723, "Ge": 72.631, "As": 74.922, "Se": 78.971, "Br": 79.904, "Kr": 84.798,
    "Rb": 84.468, "Sr": 87.62, "Y": 88.906, "Zr": 91.224, "Nb": 92.906, "Mo": 95.95,
    "Tc": 98.907, "Ru": 101.07, "Rh": 102.906, "Pd": 106.42, "Ag": 107.868, "Cd": 112.414,
    "In": 114.818, "Sn": 118.711, "Sb": 121.760, "Te": 126.7, "I": 126.904, "Xe": 131.294,
    "Cs": 132.905, "Ba": 137.328, "La": 138.905, "Ce": 140.116, "Pr": 140.908,
    "Nd": 144.243, "Pm": 144.913, "Sm": 150.36, "Eu": 151.964, "Gd": 157.25,
    "Tb": 158.925, "Dy": 162.500, "Ho": 164.930, "Er": 167.259, "Tm": 168.934,
    "Yb": 173.055, "Lu": 174.967, "Hf": 178.49, "Ta": 180.948, "W": 183.84,
    "Re": 186.207, "Os": 190.23, "Ir": 192.217, "Pt": 195.085, "Au": 196.967,
    "Hg": 200.592, "Tl": 204.383, "Pb": 207.2, "Bi": 208.980, "Po": 208.982,
    "At": 209.987, "Rn": 222.081,
}

Element = str


@dataclass
class ElementMap:
    """Element <-> 1-based atom type (ordered by atomic number) <-> mass.

    Mirrors atoms/element.py:73-150: ``unique_elements`` is sorted by name,
    atom types by atomic number.
    """

    unique_elements: Tuple[Element, ...]
    element_to_atomic_number: Dict[Element, int]
    element_to_atom_type: Dict[Element, int]
    atom_type_to_element: Dict[int, Element]

    @classmethod
    def from_list(cls, elements: Sequence[Element]) -> "ElementMap":
        unique = tuple(sorted(set(elements)))
        for e in unique:
            if e not in _ATOMIC_NUMBER:
                raise KeyError(f"unknown element '{e}'")
        z = {e: _ATOMIC_NUMBER[e] for e in unique}
        to_type = {e: t for t, e in enumerate(sorted(z, key=z.get), start=1)}
        return cls(unique, z, to_type, {t: e for e, t in to_type.items()})

    def get_atom_type_from_element(self, name: Element) -> int:
        return self.element_to_atom_type[name]

    def get_element_from_atom_type(self, value: int) -> Element:
        return self.atom_type_to_element[int(value)]

    @staticmethod
    def get_element_from_atomic_number(value: int) -> Element:
        return _ELEMENTS[value - 1]

    @staticmethod
    def get_atomic_number_from_element(name: Element) -> int:
        return _ATOMIC_NUMBER[name]

    @staticmethod
    def get_atomic_mass_from_element(name: Element) -> float:
        return _WEIGHT_U[name] * units.FROM_ATOMIC_MASS

    @property
    def n_types(self) -> int:
        return len(self.element_to_atom_type)

    def elements_by_type(self) -> Tuple[Element, ...]:
        return tuple(self.atom_type_to_element[t] for t in range(1, self.n_types + 1))

    def get_masses_from_structure(self, structure: "Structure"):
        """Masses (N, 1) FP64 on the device (element.py:135-147)."""
        import torch

        table = np.array([0.0] + [self.get_atomic_mass_from_element(e) for e in self.elements_by_type()])
        m = table[structure.atom_types_host]
        return torch.as_tensor(m, dtype=torch.float64, device=structure.device).reshape(-1, 1)

    def __repr__(self) -> str:
        return f"ElementMap(element_to_atom_type={self.element_to_atom_type})"


@dataclass
class Box:
    """Orthogonal periodic box (atoms/box.py:14-109)."""

    lattice: Any  # (3, 3) float64 device tensor

    @classmethod
    def from_list(cls, data: Sequence, device=None) -> "Box":
        import torch

        dev = device if device is not None else _lib.require_device()
        lat = torch.as_tensor(np.asarray(data, dtype=np.float64).reshape(3, 3), device=dev)
        return cls(lat)

    @property
    def length_host(self) -> Tuple[float, float, float]:
        d = np.diag(self.lattice.detach().cpu().numpy())
        return float(d[0]), float(d[1]), float(d[2])

    @property
    def lx(self) -> float:
        return self.length_host[0]

    @property
    def ly(self) -> float:
        return self.length_host[1]

    @property
    def lz(self) -> float:
        return self.length_host[2]

    @property
    def length(self):
        return self.lattice.diagonal()

    @property
    def volume(self) -> float:
        lx, ly, lz = self.length_host
        return lx * ly * lz

    def wrap_into_box(self, positions) -> Any:
        """jnp.remainder(positions, box) on the device (box.py:60-73,123-126)."""
        out = positions.to(dtype=_f64()).contiguous().clone()
        arr = host_box(self.length_host)
        _lib.check(_lib.load().pt_wrap_positions(out.shape[0], out.data_ptr(),
                                                 ctypes.cast(arr, ctypes.c_void_p),
                                                 _lib.stream_ptr()), "pt_wrap_positions")
        return out

    def __repr__(self) -> str:
        return f"Box(lengths={self.length_host})"


def _f64():
    import torch

    return torch.float64


def host_box(lengths: Optional[Sequence[float]]):
    """ctypes double[3] for the (host) box argument of the C ABI, or None."""
    if lengths is None:
        return None
    return (ctypes.c_double * 3)(*[float(x) for x in lengths])


_ATOM_ATTRIBUTES = ("positions", "forces", "energies", "charges", "total_energy", "total_charge")


@dataclass
class Structure:
    """Atoms of one snapshot (atoms/structure.py:21-343), stored on the GPU."""

    positions: Any
    forces: Any
    energies: Any
    charges: Any
    total_energy: Any
    total_charge: Any
    atom_types: Any
    element_map: ElementMap
    box: Optional[Box] = None
    atom_types_host: np.ndarray = field(default=None, repr=False)

    def __post_init__(self) -> None:
        if self.atom_types_host is None:
            self.atom_types_host = self.atom_types.detach().cpu().numpy().astype(np.int32)
        if self.box is not None:
            self.positions = self.box.wrap_into_box(self.positions).to(self.positions.dtype)

    # ---- construction ----------------------------------------------------
    @classmethod
    def from_dict(cls, data: Dict[str, Any], dtype=None, device=None) -> "Structure":
        """Build from lists/arrays of positions, elements, lattice, ... (structure.py:139-172)."""
        import torch

        dev = device if device is not None else _lib.require_device()
        dtype = dtype if dtype is not None else torch.float64
        d = defaultdict(list, data)
        if "positions" not in data or "elements" not in data:
            raise KeyError("Cannot find at least one of the expected keyword in the input data dictionary.")
        elements = list(d["elements"])
        emap = ElementMap.from_list(elements)
        n = len(elements)
        types = np.array([emap.element_to_atom_type[e] for e in elements], dtype=np.int32)

        def arr(key, shape):
            v = d[key]
            if v is None or (hasattr(v, "__len__") and len(v) == 0):
                return torch.zeros(shape, dtype=dtype, device=dev)
            return torch.as_tensor(np.asarray(v, dtype=np.float64).reshape(shape), dtype=dtype, device=dev)

        lattice = d["lattice"]
        box = Box.from_list(lattice, device=dev) if len(lattice) > 0 else None
        te = d["total_energy"]
        tc = d["total_charge"]
        return cls(
            positions=arr("positions", (n, 3)),
            forces=arr("forces", (n, 3)),
            energies=arr("energies", (n,)),
            charges=arr("charges", (n,)),
            total_energy=torch.as_tensor(float(np.sum(te)) if len(te) else 0.0, dtype=dtype, device=dev),
            total_charge=torch.as_tensor(float(np.sum(tc)) if len(tc) else 0.0, dtype=dtype, device=dev),
            atom_types=torch.as_tensor(types, device=dev),
            element_map=emap,
            box=box,
            atom_types_host=types,
        )

    # ---- attributes ------------------------------------------------------
    @property
    def natoms(self) -> int:
        return int(self.positions.shape[0])

    @property
    def dtype(self):
        return self.positions.dtype

    @property
    def device(self):
        return self.positions.device

    @property
    def lattice(self):
        return self.box.lattice if self.box is not None else None

    @property
    def box_lengths(self) -> Optional[Tuple[float, float, float]]:
        return self.box.length_host if self.box is not None else None

    def get_unique_elements(self) -> Tuple[Element, ...]:
        return self.element_map.unique_elements

    def get_elements(self) -> Tuple[Element, ...]:
        to_e = self.element_map.atom_type_to_element
        return tuple(to_e[int(t)] for t in self.atom_types_host)

    def select_host(self, element: Element) -> np.ndarray:
        t = self.element_map.element_to_atom_type[element]
        return np.nonzero(self.atom_types_host == t)[0]

    def select(self, element: Element):
        """Indices of the atoms of ``element``, ascending (structure.py:252-261)."""
        import torch

        return torch.as_tensor(self.select_host(element), dtype=torch.int64, device=self.device)

    def type_counts(self) -> List[int]:
        return [int(np.count_nonzero(self.atom_types_host == t)) for t in range(1, self.element_map.n_types + 1)]

    def get_positions_per_element(self) -> Dict[Element, Any]:
        return {e: self.positions[self.select(e)] for e in self.get_unique_elements()}

    def get_forces_per_element(self) -> Dict[Element, Any]:
        return {e: self.forces[self.select(e)] for e in self.get_unique_elements()}

    def to_dict(self) -> Dict[str, np.ndarray]:
        out = {a: getattr(self, a).detach().cpu().numpy() for a in _ATOM_ATTRIBUTES}
        out["atom_types"] = self.atom_types_host.copy()
        out["lattice"] = self.lattice.detach().cpu().numpy() if self.box is not None else []
        out["elements"] = list(self.get_elements())
        return out

    def copy(self) -> "Structure":
        return Structure(
            positions=self.positions.clone(), forces=self.forces.clone(), energies=self.energies.clone(),
            charges=self.charges.clone(), total_energy=self.total_energy.clone(),
            total_charge=self.total_charge.clone(), atom_types=self.atom_types.clone(),
            element_map=self.element_map, box=self.box, atom_types_host=self.atom_types_host.copy(),
        )

    def __repr__(self) -> str:
        return f"Structure(natoms={self.natoms}, elements={self.get_unique_elements()}, dtype={self.dtype})"
