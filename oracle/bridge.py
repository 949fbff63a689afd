"""Turn a pantea_amd potential/descriptor into oracle inputs -- TEST
INFRASTRUCTURE ONLY (imported by tests/, smoke() and bench.py's
cpu_baseline leg).  It reads parameters only; every number it returns comes
from the oracle (oracle.c + pyoracle.py).
"""

from __future__ import annotations

from typing import Dict, List

import numpy as np

from . import pyoracle


def sf_rows(acsf, element_to_atom_type: Dict[str, int]) -> np.ndarray:
    """Oracle rows (10 doubles) of an ACSF descriptor, in column order."""
    codes = pyoracle.CUTOFF_CODES
    rows: List[List[float]] = []
    for sf, ne in acsf.radial_symmetry_functions:
        rows.append([sf.kind, element_to_atom_type.get(ne.neighbor_j, 0), 0,
                     getattr(sf, "eta", 0.0), getattr(sf, "r_shift", 0.0), 0.0, 0.0,
                     sf.r_cutoff, codes[sf.cfn.cutoff_type], 0.0])
    for sf, ne in acsf.angular_symmetry_functions:
        rows.append([sf.kind, element_to_atom_type.get(ne.neighbor_j, 0),
                     element_to_atom_type.get(ne.neighbor_k, 0), sf.eta, sf.r_shift, sf.zeta,
                     sf.lambda0, sf.r_cutoff, codes[sf.cfn.cutoff_type], 0.0])
    return np.asarray(rows, dtype=np.float64).reshape(-1, 10)


def oracle_nnp(pot, structure) -> pyoracle.OracleNNP:
    emap = structure.element_map.element_to_atom_type
    tables = {e: sf_rows(pot.atomic_potentials[e].descriptor, emap) for e in pot.elements}
    scal = {}
    for e in pot.elements:
        p = pot.scalers_params[e]
        scal[e] = {"mean": np.asarray(p.mean), "sigma": np.asarray(p.sigma),
                   "minval": np.asarray(p.minval), "maxval": np.asarray(p.maxval)}
    layers = {}
    for e in pot.elements:
        model = pot.atomic_potentials[e].model
        acts = [a for _, a in model.hidden_layers] + [model.output_layer[1]]
        prm = pot.models_params[e]
        keys = sorted(prm, key=lambda k: int(k.split("_")[1]))
        layers[e] = [(np.asarray(prm[k]["kernel"], np.float64), np.asarray(prm[k]["bias"], np.float64), a)
                     for k, a in zip(keys, acts)]
    sc = pot.atomic_potentials[pot.elements[0]].scaler
    return pyoracle.OracleNNP(list(pot.elements), {e: emap[e] for e in pot.elements if e in emap},
                              tables, sc.scale_type, scal, layers, sc.scale_min, sc.scale_max)
