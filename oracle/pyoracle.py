"""FP64 CPU oracle for pantea's NNP path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module, and only as the checker.  The product package
(``pantea_amd``) never imports it.

The heavy geometry (dense MIC neighbour masks, ACSF G1/G2/G3/G9 values,
their analytic self-Jacobian and force vector-Jacobian products) lives in
``oracle/oracle.c`` (built into ``oracle/liboracle.so`` by ``oracle/Makefile``).
This module adds the small dense parts in numpy FP64, each restating the
reference:

  scaler transforms      descriptors/scaler.py:206-246
  scaler fit/partial_fit descriptors/scaler.py:249-283
  MLP forward/backward   models/nn/model.py:24-58, models/nn/activation.py:7-60
  total energy           potentials/nnp/energy.py:45-63 (per element, summed)
  forces                 potentials/nnp/force.py:16-43 (self term), and -grad E
  velocity Verlet        simulation/molecular_dynamics.py:16-30, :65-77
  kinetic energy, T, P   simulation/system.py:20-44, :115-140
  Berendsen rescale      simulation/thermostat.py:12-22
  masses                 atoms/element.py:28-58, :135-147; units.py:64-74

Symmetry-function rows are 10 doubles:
  [kind, ej, ek, eta, rs, zeta, lambda, rc, cutoff_code, 0]
with kind 1/2/3/9 (G1/G2/G3/G9), 1-based atom types, and the cutoff codes
0 hard, 1 cos, 2 tanhu, 3 tanh, 4 exp, 5 poly1, 6 poly2.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liboracle.so"

CUTOFF_CODES = {"hard": 0, "cos": 1, "tanhu": 2, "tanh": 3, "exp": 4, "poly1": 5, "poly2": 6}

# units.py:64-74 (Hartree atomic units) and element.py:28-58 (masses in u).
BOLTZMANN = 3.166811563e-6
TO_ATOMIC_MASS = 5.48579957163e-4
TO_FEMTO_SECOND = 2.418884326e-05 * 1000
MASS_U = {"H": 1.008, "He": 4.003, "O": 15.999, "Ne": 20.180}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_longlong)


def _load() -> ctypes.CDLL:
    if not _LIB_PATH.exists():
        raise RuntimeError(
            f"oracle library missing: {_LIB_PATH} (build it with `make -C oracle`)"
        )
    lib = ctypes.CDLL(str(_LIB_PATH))
    lib.or_neighbors.restype = ctypes.c_long
    lib.or_neighbors.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_double, _ip, _ip, ctypes.c_long]
    lib.or_acsf.restype = ctypes.c_int
    lib.or_acsf.argtypes = [ctypes.c_int, _dp, _ip, _dp, ctypes.c_int, _dp, ctypes.c_int, _dp, _ip, _dp]
    lib.or_acsf_selfjac.restype = ctypes.c_int
    lib.or_acsf_selfjac.argtypes = lib.or_acsf.argtypes
    lib.or_cell_neighbors.restype = ctypes.c_long
    lib.or_cell_neighbors.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_double, _lp, _ip, ctypes.c_long]
    lib.or_celllist_centres.restype = ctypes.c_int
    lib.or_celllist_centres.argtypes = [ctypes.c_int, _dp, _ip, _dp, _lp, _ip, ctypes.c_int, _dp, ctypes.c_int,
                                        _ip, _dp, _dp, _dp]
    lib.or_reference_algorithm.restype = ctypes.c_int
    lib.or_reference_algorithm.argtypes = [ctypes.c_int, _dp, _ip, _dp, ctypes.c_int, _dp, ctypes.c_int, _ip,
                                           _dp, _dp, _dp]
    lib.or_nnp_forces.restype = ctypes.c_int
    lib.or_nnp_forces.argtypes = [
        ctypes.c_int, _dp, _ip, _dp, ctypes.c_int, _ip, _ip, _dp,
        ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_ubyte), _dp, _dp,
    ]
    for name in ("or_cutoff", "or_cutoff_der"):
        getattr(lib, name).restype = ctypes.c_double
        getattr(lib, name).argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
    lib.or_mic.restype = ctypes.c_double
    lib.or_mic.argtypes = [ctypes.c_double, ctypes.c_double]
    lib.or_wrap.restype = ctypes.c_double
    lib.or_wrap.argtypes = [ctypes.c_double, ctypes.c_double]
    lib.or_tanh_prefactor.restype = ctypes.c_double
    lib.or_tanh_prefactor.argtypes = []
    lib.or_distances_from.restype = None
    lib.or_distances_from.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
    lib.or_wrap_positions.restype = None
    lib.or_wrap_positions.argtypes = [ctypes.c_int, _dp, _dp]
    return lib


_lib: Optional[ctypes.CDLL] = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _p(a: np.ndarray):
    if a is None:
        return None
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    if a.dtype == np.int64:
        return a.ctypes.data_as(_lp)
    if a.dtype == np.uint8:
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))
    raise TypeError(a.dtype)


def _box(box) -> Optional[np.ndarray]:
    if box is None:
        return None
    b = np.asarray(box, dtype=np.float64)
    if b.shape == (3, 3):
        b = np.diag(b)
    return _f64(b)


# ---------------------------------------------------------------------------
# geometry (oracle.c)
# ---------------------------------------------------------------------------

def cutoff(code: int, rc: float, r: float) -> float:
    return lib().or_cutoff(code, rc, r)


def cutoff_der(code: int, rc: float, r: float) -> float:
    return lib().or_cutoff_der(code, rc, r)


def tanh_prefactor() -> float:
    return lib().or_tanh_prefactor()


def mic(dx: float, L: float) -> float:
    return lib().or_mic(dx, L)


def wrap(x: float, L: float) -> float:
    return lib().or_wrap(x, L)


def wrap_positions(pos, box) -> np.ndarray:
    p = _f64(pos).copy()
    lib().or_wrap_positions(p.shape[0], _p(p), _p(_box(box)))
    return p


def distances_from(pos, box, point) -> np.ndarray:
    pos = _f64(pos)
    out = np.empty(pos.shape[0])
    lib().or_distances_from(pos.shape[0], _p(pos), _p(_box(box)), _p(_f64(point)), _p(out))
    return out


def neighbors(pos, box, rc: float) -> Tuple[np.ndarray, np.ndarray]:
    """Dense MIC masks as CSR (counts, idx), rows ascending (neighbor.py:102-107)."""
    pos = _f64(pos)
    n = pos.shape[0]
    counts = np.zeros(n, dtype=np.int32)
    cap = max(1024, n * 256)
    while True:
        idx = np.zeros(cap, dtype=np.int32)
        tot = lib().or_neighbors(n, _p(pos), _p(_box(box)), rc, _p(counts), _p(idx), cap)
        if tot >= 0:
            return counts, idx[:tot].copy()
        cap = -tot


def acsf(pos, types, box, sf_rows, center_pos=None, self_idx=None) -> np.ndarray:
    """Descriptor rows (acsf.py:163-202) for centres at ``center_pos``."""
    pos = _f64(pos)
    types = _i32(types)
    sf = _f64(np.asarray(sf_rows, dtype=np.float64).reshape(-1, 10))
    if center_pos is None:
        center_pos = pos
        self_idx = np.arange(pos.shape[0], dtype=np.int32)
    cpos = _f64(center_pos).reshape(-1, 3)
    sidx = None if self_idx is None else _i32(self_idx)
    out = np.zeros((cpos.shape[0], sf.shape[0]))
    rc = lib().or_acsf(pos.shape[0], _p(pos), _p(types), _p(_box(box)), sf.shape[0], _p(sf),
                       cpos.shape[0], _p(cpos), _p(sidx), _p(out))
    if rc != 0:
        raise MemoryError("oracle acsf failed")
    return out


def cell_neighbors(pos, box, rc: float) -> Tuple[np.ndarray, np.ndarray]:
    """(row_off, idx): the rows of neighbors() built with a cell list (O(N))."""
    pos = _f64(pos)
    n = pos.shape[0]
    row_off = np.zeros(n + 1, dtype=np.int64)
    cap = max(1024, 200 * n)
    while True:
        idx = np.zeros(cap, dtype=np.int32)
        t = lib().or_cell_neighbors(n, _p(pos), _p(_box(box)), float(rc), _p(row_off), _p(idx), cap)
        if t >= 0:
            return row_off, idx[:t]
        cap = -t + 1024


def celllist_centres(pos, types, box, row_off, idx, sf_rows, centres, w=None):
    """Descriptors (and with w the self gradient sum_s w_s dG_is/dr_i) of ``centres`` from CSR rows."""
    pos = _f64(pos)
    types = _i32(types)
    sf = _f64(np.asarray(sf_rows, dtype=np.float64).reshape(-1, 10))
    cen = _i32(centres)
    G = np.zeros((cen.size, sf.shape[0]))
    grad = np.zeros((cen.size, 3)) if w is not None else None
    wp = None if w is None else _f64(np.asarray(w, dtype=np.float64).reshape(cen.size, sf.shape[0]))
    rc = lib().or_celllist_centres(pos.shape[0], _p(pos), _p(types), _p(_box(box)),
                                   _p(np.ascontiguousarray(row_off, dtype=np.int64)),
                                   _p(np.ascontiguousarray(idx, dtype=np.int32)), sf.shape[0], _p(sf), cen.size,
                                   _p(cen), _p(G), _p(wp), _p(grad))
    if rc != 0:
        raise MemoryError("oracle cell-list centres failed")
    return G, grad


def reference_algorithm(pos, types, box, sf_rows, centres, w=None):
    """Descriptors (and, with w, the gradient of sum_s w_s G_is w.r.t. each
    centre) by the reference's dense algorithm: all N atoms per radial
    function and all N x N ordered pairs per angular function, every term
    evaluated and masked (acsf.py:163-330).  Used to time the CPU baseline."""
    pos = _f64(pos)
    types = _i32(types)
    sf = _f64(np.asarray(sf_rows, dtype=np.float64).reshape(-1, 10))
    cen = _i32(centres)
    G = np.zeros((cen.size, sf.shape[0]))
    grad = np.zeros((cen.size, 3)) if w is not None else None
    wp = None if w is None else _f64(np.asarray(w, dtype=np.float64).reshape(cen.size, sf.shape[0]))
    rc = lib().or_reference_algorithm(pos.shape[0], _p(pos), _p(types), _p(_box(box)), sf.shape[0], _p(sf),
                                      cen.size, _p(cen), _p(G), _p(wp), _p(grad))
    if rc != 0:
        raise MemoryError("oracle reference algorithm failed")
    return G, grad


def acsf_selfjac(pos, types, box, sf_rows, center_pos=None, self_idx=None) -> np.ndarray:
    """dG_i/dr_i, shape (n_centres, n_sf, 3) (acsf.py:215-228 semantics)."""
    pos = _f64(pos)
    types = _i32(types)
    sf = _f64(np.asarray(sf_rows, dtype=np.float64).reshape(-1, 10))
    if center_pos is None:
        center_pos = pos
        self_idx = np.arange(pos.shape[0], dtype=np.int32)
    cpos = _f64(center_pos).reshape(-1, 3)
    sidx = None if self_idx is None else _i32(self_idx)
    out = np.zeros((cpos.shape[0], sf.shape[0], 3))
    rc = lib().or_acsf_selfjac(pos.shape[0], _p(pos), _p(types), _p(_box(box)), sf.shape[0],
                               _p(sf), cpos.shape[0], _p(cpos), _p(sidx), _p(out))
    if rc != 0:
        raise MemoryError("oracle selfjac failed")
    return out


def force_vjp(pos, types, box, tables: Sequence[np.ndarray], w: np.ndarray,
              want_center: Optional[np.ndarray] = None) -> Tuple[np.ndarray, np.ndarray]:
    """f_self, f_full from per-atom weights w[i, s] = dE_i/dG_is.

    ``tables[t-1]`` is the SF table of atom type t; w has one row per atom,
    padded to the widest table.
    """
    pos = _f64(pos)
    types = _i32(types)
    n = pos.shape[0]
    offs, cnts, rows = [], [], []
    o = 0
    for t in tables:
        t = np.asarray(t, dtype=np.float64).reshape(-1, 10)
        offs.append(o)
        cnts.append(t.shape[0])
        rows.append(t)
        o += t.shape[0]
    sf_all = _f64(np.concatenate(rows, axis=0) if rows else np.zeros((0, 10)))
    w = _f64(w)
    f_self = np.zeros((n, 3))
    f_full = np.zeros((n, 3))
    wc = None if want_center is None else np.ascontiguousarray(want_center, dtype=np.uint8)
    rc = lib().or_nnp_forces(n, _p(pos), _p(types), _p(_box(box)), len(tables),
                             _p(_i32(offs)), _p(_i32(cnts)), _p(sf_all), w.shape[1], _p(w),
                             _p(wc), _p(f_self), _p(f_full))
    if rc != 0:
        raise MemoryError("oracle force vjp failed")
    return f_self, f_full


# ---------------------------------------------------------------------------
# scaler (scaler.py:206-283)
# ---------------------------------------------------------------------------

def scaler_transform(kind: str, params: Dict[str, np.ndarray], x: np.ndarray,
                     smin: float = 0.0, smax: float = 1.0) -> np.ndarray:
    x = np.atleast_2d(x)
    if kind == "center":
        return x - params["mean"]
    if kind == "scale":
        return smin + (smax - smin) * (x - params["minval"]) / (params["maxval"] - params["minval"])
    if kind == "scale_center":
        return smin + (smax - smin) * (x - params["mean"]) / (params["maxval"] - params["minval"])
    if kind == "scale_center_sigma":
        return smin + (smin - smax) * (x - params["mean"]) / params["sigma"]
    raise ValueError(kind)


def scaler_slope(kind: str, params: Dict[str, np.ndarray], smin: float = 0.0,
                 smax: float = 1.0) -> np.ndarray:
    """d(scaled)/d(x), per feature."""
    if kind == "center":
        return np.ones_like(np.asarray(params["mean"], dtype=np.float64))
    if kind in ("scale", "scale_center"):
        return (smax - smin) / (params["maxval"] - params["minval"])
    if kind == "scale_center_sigma":
        return (smin - smax) / params["sigma"]
    raise ValueError(kind)


def scaler_fit(x: np.ndarray) -> Dict[str, np.ndarray]:
    x = np.atleast_2d(np.asarray(x, dtype=np.float64))
    return {"dimension": x.shape[1], "nsamples": x.shape[0], "mean": x.mean(0),
            "sigma": x.std(0), "minval": x.min(0), "maxval": x.max(0)}


def scaler_partial_fit(params: Dict[str, np.ndarray], x: np.ndarray) -> Dict[str, np.ndarray]:
    x = np.atleast_2d(np.asarray(x, dtype=np.float64))
    m, n = params["nsamples"], x.shape[0]
    new_mean, new_sigma = x.mean(0), x.std(0)
    fm, fn = m / (m + n), n / (m + n)
    diff = params["mean"] - new_mean
    return {
        "dimension": params["dimension"],
        "nsamples": m + n,
        "mean": fm * params["mean"] + fn * new_mean,
        "sigma": np.sqrt((fm * params["sigma"]) * params["sigma"]
                         + (fn * new_sigma) * new_sigma + (fm * diff) * (fn * diff)),
        "minval": np.minimum(params["minval"], x.min(0)),
        "maxval": np.maximum(params["maxval"], x.max(0)),
    }


# ---------------------------------------------------------------------------
# MLP (model.py:24-58, activation.py:7-60)
# ---------------------------------------------------------------------------

def _act(name: str, z: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    if name == "identity":
        return z, np.ones_like(z)
    if name == "tanh":
        t = np.tanh(z)
        return t, 1.0 - t * t
    if name == "logistic":
        s = 1.0 / (1.0 + np.exp(-z))
        return s, s * (1.0 - s)
    if name == "softplus":
        return np.logaddexp(z, 0.0), 1.0 / (1.0 + np.exp(-z))
    if name == "relu":
        return np.maximum(z, 0.0), (z > 0).astype(z.dtype)
    if name == "gaussian":
        g = np.exp(-0.5 * z * z)
        return g, -z * g
    if name == "cos":
        return np.cos(z), -np.sin(z)
    if name == "exp":
        e = np.exp(-z)
        return e, -e
    if name == "harmonic":
        return z * z, 2.0 * z
    raise ValueError(name)


def mlp_forward_backward(layers: List[Tuple[np.ndarray, np.ndarray, str]],
                         x: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """y = MLP(x) and dy/dx for a scalar-output MLP; layers = [(W(in,out), b, act)]."""
    h = np.asarray(x, dtype=np.float64)
    derivs = []
    for W, b, a in layers:
        z = h @ np.asarray(W, dtype=np.float64) + np.asarray(b, dtype=np.float64)
        h, d = _act(a, z)
        derivs.append(d)
    y = h[:, 0]
    g = derivs[-1]  # (n, 1)
    for li in range(len(layers) - 1, -1, -1):
        W = np.asarray(layers[li][0], dtype=np.float64)
        g = g @ W.T
        if li > 0:
            g = g * derivs[li - 1]
    return y, g


# ---------------------------------------------------------------------------
# NNP energy and forces (energy.py:23-63, force.py:16-43, potential.py:83-102)
# ---------------------------------------------------------------------------

class OracleNNP:
    """Per-element (SF table, scaler, MLP) bundle evaluated in FP64.

    ``elements``: element names in potential order; ``types``: 1-based atom
    type of each element in the structure's element map.
    """

    def __init__(self, elements: Sequence[str], type_of: Dict[str, int],
                 tables: Dict[str, np.ndarray], scaler_kind: str,
                 scaler_params: Dict[str, Dict[str, np.ndarray]],
                 layers: Dict[str, List[Tuple[np.ndarray, np.ndarray, str]]],
                 smin: float = 0.0, smax: float = 1.0):
        self.elements = list(elements)
        self.type_of = dict(type_of)
        self.tables = {e: np.asarray(tables[e], dtype=np.float64).reshape(-1, 10) for e in elements}
        self.scaler_kind = scaler_kind
        self.scaler_params = scaler_params
        self.layers = layers
        self.smin, self.smax = smin, smax

    def _tables_by_type(self) -> List[np.ndarray]:
        nt = max(self.type_of.values())
        out = [np.zeros((0, 10)) for _ in range(nt)]
        for e in self.elements:
            out[self.type_of[e] - 1] = self.tables[e]
        return out

    def descriptors(self, pos, types, box) -> Dict[str, np.ndarray]:
        out = {}
        for e in self.elements:
            idx = np.nonzero(np.asarray(types) == self.type_of[e])[0]
            out[e] = acsf(pos, types, box, self.tables[e], center_pos=np.asarray(pos)[idx],
                          self_idx=idx.astype(np.int32))
        return out

    def atomic(self, pos, types, box):
        """Per-atom energies and w = dE_i/dG_i (rows padded to the widest table)."""
        n = len(types)
        width = max(t.shape[0] for t in self.tables.values())
        e_atom = np.zeros(n)
        w = np.zeros((n, width))
        G = self.descriptors(pos, types, box)
        for e in self.elements:
            idx = np.nonzero(np.asarray(types) == self.type_of[e])[0]
            if idx.size == 0:
                continue
            x = scaler_transform(self.scaler_kind, self.scaler_params[e], G[e], self.smin, self.smax)
            y, dy = mlp_forward_backward(self.layers[e], x)
            slope = scaler_slope(self.scaler_kind, self.scaler_params[e], self.smin, self.smax)
            e_atom[idx] = y
            w[idx, : dy.shape[1]] = dy * slope
        return e_atom, w, G

    def energy(self, pos, types, box) -> float:
        e_atom, _, _ = self.atomic(pos, types, box)
        total = 0.0
        for e in self.elements:
            idx = np.nonzero(np.asarray(types) == self.type_of[e])[0]
            total += e_atom[idx].sum()
        return float(total)

    def energy_forces(self, pos, types, box, want_center=None):
        e_atom, w, _ = self.atomic(pos, types, box)
        total = 0.0
        for e in self.elements:
            idx = np.nonzero(np.asarray(types) == self.type_of[e])[0]
            total += e_atom[idx].sum()
        f_self, f_full = force_vjp(pos, types, box, self._tables_by_type(), w, want_center)
        return float(total), e_atom, f_self, f_full


# ---------------------------------------------------------------------------
# MD pieces (molecular_dynamics.py:16-30, system.py:20-44, thermostat.py:12-22)
# ---------------------------------------------------------------------------

def mass_au(element: str) -> float:
    return MASS_U[element] / TO_ATOMIC_MASS


def verlet_positions(x, v, a, dt, box=None):
    xn = x + v * dt + 0.5 * a * dt * dt
    if box is not None:
        L = np.diag(box) if np.ndim(box) == 2 else np.asarray(box)
        xn = np.where(np.fmod(xn, L) != 0, np.fmod(xn, L), 0.0)
        xn = np.where(xn < 0, xn + L, xn)
    return xn


def verlet_velocities(v, a_old, a_new, dt):
    return v + 0.5 * (a_old + a_new) * dt


def kinetic_energy(v, m) -> float:
    return float(0.5 * np.sum(np.asarray(m).reshape(-1, 1) * v * v))


def temperature(v, m) -> float:
    return 2.0 * kinetic_energy(v, m) / (3 * v.shape[0] * BOLTZMANN)


def berendsen_factor(dt, tau, T, T0) -> float:
    return 1.0 / np.sqrt(1.0 + (dt / tau) * (T / T0 - 1.0))
