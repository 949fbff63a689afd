"""Synthetic liquid-water structures for the benchmark and tests.

This generator has no counterpart in the reference.  Recipe:
  * O positions by random sequential addition, uniform in [0, L)^3, rejected
    when the minimum-image O-O distance to an accepted O is below d_min;
  * each H at r_OH from its O with H-O-H angle theta, orientation uniformly
    random (random unit bisector, random rotation about it);
  * Gaussian jitter of sigma on every atom, then wrap into [0, L)^3;
  * atom order O, H, H per molecule.
Randomness comes from numpy's PCG64 seeded with ``seed``.  Lengths in Bohr.
"""

from __future__ import annotations

import math
from typing import Dict, Tuple

import numpy as np

from .units import units

AVOGADRO = 6.02214076e23
WATER_G_PER_MOL = 18.015

# Benchmark shapes: (molecules, box length in Bohr).
CONFIGS: Dict[str, Tuple[int, float]] = {
    "cfg0": (128, 29.56),
    "cfg1": (4096, 93.85),
    "cfg2": (32768, 187.7),
}


def box_for_density(n_molecules: int, density_g_cm3: float = 1.0) -> float:
    """Cubic box length (Bohr) holding n water molecules at the given density."""
    vol_cm3 = n_molecules * WATER_G_PER_MOL / AVOGADRO / density_g_cm3
    return (vol_cm3 ** (1.0 / 3.0)) * 1e8 * units.FROM_ANGSTROM


def _place_oxygens(n: int, L: float, d_min: float, rng: np.random.Generator) -> np.ndarray:
    nc = max(1, int(L // d_min))
    w = L / nc
    grid: Dict[Tuple[int, int, int], list] = {}
    out = np.empty((n, 3))
    d2 = d_min * d_min
    count = 0
    attempts = 0
    max_attempts = 400 * n + 10000
    while count < n:
        attempts += 1
        if attempts > max_attempts:
            raise RuntimeError(f"could not place {n} O atoms with d_min={d_min} in L={L}")
        p = rng.random(3) * L
        c = tuple(int(x / w) % nc for x in p)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for q in grid.get(((c[0] + dx) % nc, (c[1] + dy) % nc, (c[2] + dz) % nc), ()):
                        d = p - out[q]
                        d -= L * np.round(d / L)
                        if d @ d < d2:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            out[count] = p
            grid.setdefault(c, []).append(count)
            count += 1
    return out


def water_box(n_molecules: int, L: float, seed: int = 2023, d_min: float = 4.5, r_oh: float = 1.81,
              theta_deg: float = 104.5, jitter: float = 0.05) -> Dict:
    """Positions/elements/lattice dict accepted by ``Structure.from_dict``."""
    rng = np.random.default_rng(seed)
    O = _place_oxygens(n_molecules, L, d_min, rng)
    u = rng.normal(size=(n_molecules, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    a = rng.normal(size=(n_molecules, 3))
    v = a - np.sum(a * u, axis=1, keepdims=True) * u
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    half = math.radians(theta_deg) / 2.0
    h1 = O + r_oh * (math.cos(half) * u + math.sin(half) * v)
    h2 = O + r_oh * (math.cos(half) * u - math.sin(half) * v)
    pos = np.stack([O, h1, h2], axis=1).reshape(-1, 3)
    pos = pos + rng.normal(scale=jitter, size=pos.shape)
    pos = np.mod(pos, L)
    pos[pos >= L] = 0.0
    elements = ["O", "H", "H"] * n_molecules
    return {"positions": pos, "elements": elements, "lattice": np.diag([L, L, L])}


def config(name: str, seed: int = 2023) -> Dict:
    n, L = CONFIGS[name]
    return water_box(n, L, seed=seed)
