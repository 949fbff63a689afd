"""The water NNPs of the benchmark (teacher and student).

Teacher (per central element, tanh cutoff rc = 12 Bohr):
  * 8 G2 per neighbour element: eta log-spaced in [0.001, 0.5] Bohr^-2, r_shift 0
  * 12 G3 per neighbour-element pair: eta in {0.001, 0.01, 0.05}, zeta in {1, 4},
    lambda = +1/-1
  -> 16 + 36 = 52 functions; MLP 52 -> h -> h -> 1, tanh / linear,
  weights ~ N(0, 1/fan_in), zero biases, min-max scaler fitted on a structure.
Student: 8 G2 per neighbour element and 6 G3 per pair -> 34 functions; MLP
34 -> 16 -> 16 -> 1.  Its cutoff and its 6 G3 per pair (cells of the
teacher's 12: eta x zeta x lambda) are chosen per central element and stored
with its weights (student.json); without that file it uses zeta = 1 and the
teacher's cutoff.
"""

from __future__ import annotations

import itertools
import json
from pathlib import Path
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from .descriptors import ACSF, G2, G3, CutoffFunction, DescriptorScaler, NeighborElements
from .models import NeuralNetworkModel, NormalInitializer
from .potential import AtomicPotential, NeuralNetworkPotential

ELEMENTS = ("H", "O")
PAIRS = (("H", "H"), ("H", "O"), ("O", "O"))


def teacher_acsf(center: str, r_cutoff: float = 12.0, cutoff: str = "tanh") -> ACSF:
    cfn = CutoffFunction.from_type(cutoff, r_cutoff)
    etas = np.geomspace(0.001, 0.5, 8)
    radial = tuple((G2(cfn, r_shift=0.0, eta=float(eta)), NeighborElements(e)) for e in ELEMENTS for eta in etas)
    angular = tuple(
        (G3(cfn, eta=eta, zeta=zeta, lambda0=lam, r_shift=0.0), NeighborElements(j, k))
        for (j, k) in PAIRS for eta in (0.001, 0.01, 0.05) for zeta in (1.0, 4.0) for lam in (1.0, -1.0))
    return ACSF(center, radial, angular)


# the teacher's 12 angular cells per neighbour pair: (eta, zeta, lambda)
G3_GRID = tuple((eta, zeta, lam) for eta in (0.001, 0.01, 0.05) for zeta in (1.0, 4.0) for lam in (1.0, -1.0))
ZETA1_CELLS = tuple(c for c in G3_GRID if c[1] == 1.0)


def pair_key(j: str, k: str) -> str:
    return f"{j}-{k}"


def student_acsf(center: str, r_cutoff: float = 12.0, cutoff: str = "tanh",
                 g3_cells: Optional[Dict[str, Sequence[Tuple[float, float, float]]]] = None) -> ACSF:
    """8 G2 per neighbour element and, per neighbour pair, the angular cells
    ``g3_cells["H-O"]`` etc. (default: the zeta = 1 cells, six per pair)."""
    cfn = CutoffFunction.from_type(cutoff, r_cutoff)
    etas = np.geomspace(0.001, 0.5, 8)
    radial = tuple((G2(cfn, r_shift=0.0, eta=float(eta)), NeighborElements(e)) for e in ELEMENTS for eta in etas)
    angular = []
    for (j, k) in PAIRS:
        cells = ZETA1_CELLS if g3_cells is None else g3_cells[pair_key(j, k)]
        for eta, zeta, lam in cells:
            angular.append((G3(cfn, eta=float(eta), zeta=float(zeta), lambda0=float(lam), r_shift=0.0),
                            NeighborElements(j, k)))
    return ACSF(center, radial, tuple(angular))


def make_nnp(acsf_of, hidden: Sequence[int], seed: int, precision: str = "float64",
             scale_type: str = "scale") -> NeuralNetworkPotential:
    aps = {}
    params = {}
    for i, e in enumerate(ELEMENTS):
        model = NeuralNetworkModel(hidden_layers=tuple((h, "tanh") for h in hidden),
                                   output_layer=(1, "identity"), kernel_initializer=NormalInitializer(1.0))
        acsf = acsf_of(e)
        aps[e] = AtomicPotential(acsf, DescriptorScaler.from_type(scale_type), model)
        params[e] = model.init(seed + 101 * (i + 1), acsf.num_symmetry_functions)
    return NeuralNetworkPotential(directory=".", elements=ELEMENTS, atomic_potentials=aps, models_params=params,
                                  scalers_params={e: None for e in ELEMENTS}, precision=precision)


def water_nnp(structure, hidden: Sequence[int] = (25, 25), seed: int = 2023, precision: str = "float64",
              student: bool = False) -> NeuralNetworkPotential:
    """Teacher (or student) NNP with its min-max scaler fitted on ``structure``."""
    pot = make_nnp(student_acsf if student else teacher_acsf, hidden, seed, precision)
    pot.fit_scaler([structure])
    return pot


def water_teacher(precision: str = "mixed", molecules: int = 32768, box: float = 187.7, seed: int = 2023):
    """The 98,304-atom water structure and its teacher (MLP 52-64-64-1, scaler fitted on it)."""
    from . import water
    from .atoms import Structure

    s = Structure.from_dict(water.water_box(molecules, box, seed=seed))
    return s, water_nnp(s, hidden=(64, 64), seed=seed, precision=precision)


STUDENT_SPEC = "student.json"


def student_spec(r_cutoff: float = 12.0, g3: Optional[Dict[str, Dict[str, Sequence]]] = None) -> dict:
    """Descriptor choice of a student: cutoff and, per central element, its angular cells per pair."""
    g3 = g3 or {e: {pair_key(j, k): [list(c) for c in ZETA1_CELLS] for (j, k) in PAIRS} for e in ELEMENTS}
    return {"r_cutoff": float(r_cutoff), "hidden": [16, 16],
            "g3": {e: {p: [[float(x) for x in c] for c in cells] for p, cells in d.items()} for e, d in g3.items()}}


def student_from_spec(spec: dict, seed: int = 7, precision: str = "mixed") -> NeuralNetworkPotential:
    rc = float(spec["r_cutoff"])
    g3 = spec["g3"]
    return make_nnp(lambda e: student_acsf(e, r_cutoff=rc, g3_cells={p: [tuple(c) for c in cells]
                                                                     for p, cells in g3[e].items()}),
                    tuple(spec.get("hidden", (16, 16))), seed, precision)


def water_student(directory=None, seed: int = 7, precision: str = "mixed") -> NeuralNetworkPotential:
    """Student NNP (34 ACSF, MLP 34-16-16-1); with ``directory``, its descriptor choice
    (student.json), scaler and weights are loaded from there."""
    spec = student_spec()
    if directory is not None and (Path(directory) / STUDENT_SPEC).exists():
        spec = json.loads((Path(directory) / STUDENT_SPEC).read_text())
    pot = student_from_spec(spec, seed, precision)
    pot.student_spec = spec
    if directory is not None:
        pot.directory = directory
        pot.load()
    return pot
