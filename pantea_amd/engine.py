"""Device engine: neighbour lists, symmetry-function tables and the kernel
wrappers of the C ABI.  Everything here launches HIP kernels from
libpantea_hip.so on torch's current stream; torch only provides device
memory.

Internal atom order (produced by the neighbour build): atoms grouped by atom
type, then by cell, ties in input order.  ``perm[i]`` is the input index of
internal atom ``i``; type ``t`` occupies the internal block
``[type_offsets[t-1], type_offsets[t])``.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .atoms import host_box


def _torch():
    import torch

    return torch


def _vp(t) -> Optional[int]:
    return None if t is None else t.data_ptr()


def _box_arg(lengths):
    arr = host_box(lengths)
    return (ctypes.cast(arr, ctypes.c_void_p), arr) if arr is not None else (None, None)


# ---------------------------------------------------------------------------
# neighbour list
# ---------------------------------------------------------------------------

class NeighborList:
    """CSR neighbour list built on the GPU (pt_nbr_build).

    Row i (internal order) holds every j with r_ij <= r_cutoff and r_ij > 0
    under the single-shift minimum image -- the reference's cutoff mask
    (atoms/neighbor.py:102-107) -- in ascending internal index.
    """

    def __init__(self, r_cutoff: float, n_types: int):
        self.r_cutoff = float(r_cutoff)
        self.n_types = int(n_types)
        self.n = 0
        self.cap = 0
        self.ws = None
        self.nbr = None
        self.rev = None
        self.n_pairs = 0
        self.max_row = 0
        self.box_lengths = None

    def build(self, positions, types, box_lengths: Optional[Sequence[float]],
              type_counts: Sequence[int], with_rev: bool = True) -> "NeighborList":
        torch = _torch()
        lib = _lib.load()
        dev = positions.device
        n = int(positions.shape[0])
        pos = positions.to(torch.float64).contiguous()
        stride = int(pos.shape[1])
        types = types.to(torch.int32).contiguous()
        if n != self.n or self.ws is None:
            self.n = n
            self.perm = torch.empty(n, dtype=torch.int32, device=dev)
            self.pos4 = torch.empty((n, 4), dtype=torch.float64, device=dev)
            self.row_off = torch.empty(n + 1, dtype=torch.int64, device=dev)
            wsb = lib.pt_nbr_workspace_bytes(n, self.n_types)
            self.ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
            if self.cap == 0:
                self.cap = max(1024, n * 160)
                self.nbr = torch.empty(self.cap, dtype=torch.int32, device=dev)
                self.rev = torch.empty(self.cap, dtype=torch.int32, device=dev)
        self.box_lengths = None if box_lengths is None else tuple(float(x) for x in box_lengths)
        box_p, _keep = _box_arg(self.box_lengths)
        n_pairs = ctypes.c_longlong(0)
        max_row = ctypes.c_int(0)
        for _attempt in range(3):
            rc = lib.pt_nbr_build(
                n, pos.data_ptr(), stride, types.data_ptr(), self.n_types, box_p, self.r_cutoff,
                self.perm.data_ptr(), self.pos4.data_ptr(), self.row_off.data_ptr(),
                self.nbr.data_ptr(), self.rev.data_ptr() if with_rev else None, self.cap,
                ctypes.byref(n_pairs), ctypes.byref(max_row), self.ws.data_ptr(), self.ws.numel(),
                _lib.stream_ptr())
            if rc == _lib.PT_ERR_CAPACITY:
                self.cap = int(n_pairs.value * 1.25) + 1024
                self.nbr = torch.empty(self.cap, dtype=torch.int32, device=dev)
                self.rev = torch.empty(self.cap, dtype=torch.int32, device=dev)
                continue
            _lib.check(rc, "pt_nbr_build")
            break
        else:
            raise _lib.HipLibraryError(f"pt_nbr_build: {n_pairs.value} pairs still exceed the capacity "
                                       f"{self.cap} after three attempts")
        self.n_pairs = int(n_pairs.value)
        self.max_row = int(max_row.value)
        self.has_rev = with_rev
        counts = list(type_counts) + [0] * (self.n_types - len(type_counts))
        self.type_offsets = [0]
        for c in counts[: self.n_types]:
            self.type_offsets.append(self.type_offsets[-1] + int(c))
        self._perm_host = None
        return self

    def release_workspace(self) -> None:
        """Drop the build workspace and the unused capacity of a list that will not be rebuilt.

        A later build() allocates them again.
        """
        self.ws = None
        self.nbr = self.nbr[: max(self.n_pairs, 1)].clone()
        if self.rev is not None:
            self.rev = self.rev[: max(self.n_pairs, 1)].clone()
        self.cap = 0

    @property
    def perm_host(self) -> np.ndarray:
        if self._perm_host is None:
            self._perm_host = self.perm.cpu().numpy()
        return self._perm_host

    def inverse_perm_host(self) -> np.ndarray:
        inv = np.empty(self.n, dtype=np.int64)
        inv[self.perm_host] = np.arange(self.n)
        return inv

    def type_block(self, t: int) -> Tuple[int, int]:
        """(begin, count) of the internal block of atom type t (1-based)."""
        b = self.type_offsets[t - 1]
        return b, self.type_offsets[t] - b

    def to_csr_input_order(self) -> Tuple[np.ndarray, np.ndarray]:
        """(counts, idx) in input order with ascending neighbours, for checks."""
        perm = self.perm_host
        off = self.row_off.cpu().numpy()
        nbr = self.nbr[: self.n_pairs].cpu().numpy()
        counts = np.zeros(self.n, dtype=np.int32)
        rows = [None] * self.n
        for i in range(self.n):
            js = np.sort(perm[nbr[off[i]:off[i + 1]]])
            rows[perm[i]] = js
            counts[perm[i]] = js.size
        idx = np.concatenate([r for r in rows]) if self.n else np.zeros(0, np.int32)
        return counts, idx.astype(np.int32)


# ---------------------------------------------------------------------------
# symmetry-function tables
# ---------------------------------------------------------------------------

@dataclass(frozen=True)
class SFRow:
    """One symmetry function with resolved neighbour types.

    kind: 1=G1 2=G2 3=G3 9=G9; ej/ek: 1-based atom types (ek = 0 for radial);
    cut: cutoff code; col: output column.
    """

    kind: int
    ej: int
    ek: int
    eta: float
    rs: float
    zeta: float
    lam: float
    rc: float
    cut: int
    col: int

    def as_oracle_row(self) -> List[float]:
        return [self.kind, self.ej, self.ek, self.eta, self.rs, self.zeta, self.lam, self.rc, self.cut, 0.0]


TANH_PRE = ((math.e + 1 / math.e) / (math.e - 1 / math.e)) ** 3  # cutoff.py:82


_Z14_TERMS = ((1.0, 1.0), (1.0, -1.0), (4.0, 1.0), (4.0, -1.0))
_TANH_CUTS = (3, 2)   # PT_CUT_TANH, PT_CUT_TANHU


def _pad_to_z14(angular: List[dict]) -> None:
    """Complete every angular group to the four terms zeta {1, 4} x lambda +-1
    (the added terms own no column) when each group's terms are a subset of
    them under a tanh cutoff and the table is not of the cheaper class with
    at most two zeta = 1 terms per group.  The table then runs the kernels of
    the four-term class (shared angular powers, the quartic adjoint), whose
    per-pair cost is below that of the general integer-zeta class; the
    descriptor values do not change."""
    if not angular:
        return
    z1n2 = all(len(g["angs"]) <= 2 and all(z == 1.0 for z, _ in g["angs"]) for g in angular)
    subset = all(g["key"][2] in _TANH_CUTS and all(a in _Z14_TERMS for a in g["angs"]) for g in angular)
    if z1n2 or not subset:
        return
    for g in angular:
        for a in _Z14_TERMS:
            if a not in g["angs"]:
                g["angs"].append(a)


def compile_table(rows: Sequence[SFRow]) -> _lib.SFTable:
    """Group the functions into the radial/angular groups of pt_sf_table."""
    t = _lib.SFTable()
    t.n_sf = len(rows)
    t.tanh_pre = TANH_PRE
    t.rc_max = max([r.rc for r in rows], default=0.0)
    radial: List[dict] = []
    angular: List[dict] = []
    for r in rows:
        if r.kind in (1, 2):
            eta, rs = (0.0, 0.0) if r.kind == 1 else (r.eta, r.rs)
            key = (r.ej, r.cut, r.rc)
            grp = next((g for g in radial if g["key"] == key and len(g["f"]) < _lib.PT_RG_MAX), None)
            if grp is None:
                grp = {"key": key, "f": []}
                radial.append(grp)
            grp["f"].append((eta, rs, r.col))
        elif r.kind in (3, 9):
            ej, ek = min(r.ej, r.ek), max(r.ej, r.ek)
            key = (ej, ek, r.cut, r.rc, r.kind)
            ang = (r.zeta, r.lam)
            placed = False
            for g in angular:
                if g["key"] != key:
                    continue
                etas, angs, cells = g["etas"], g["angs"], g["cells"]
                e_idx = etas.index(r.eta) if r.eta in etas else (len(etas) if len(etas) < _lib.PT_AG_ETA else -1)
                a_idx = angs.index(ang) if ang in angs else (len(angs) if len(angs) < _lib.PT_AG_ANG else -1)
                if e_idx < 0 or a_idx < 0 or (e_idx, a_idx) in cells:
                    continue
                if e_idx == len(etas):
                    etas.append(r.eta)
                if a_idx == len(angs):
                    angs.append(ang)
                cells[(e_idx, a_idx)] = r.col
                placed = True
                break
            if not placed:
                angular.append({"key": key, "etas": [r.eta], "angs": [ang], "cells": {(0, 0): r.col}})
        else:
            raise ValueError(f"unknown symmetry function kind {r.kind}")
    if len(radial) > _lib.PT_MAX_GROUPS or len(angular) > _lib.PT_MAX_GROUPS:
        raise ValueError("too many symmetry-function groups for one descriptor")
    _pad_to_z14(angular)
    t.n_rg = len(radial)
    for gi, g in enumerate(radial):
        G = t.rg[gi]
        G.elem, G.cut, G.rc = g["key"][0], g["key"][1], g["key"][2]
        G.n = len(g["f"])
        for q, (eta, rs, col) in enumerate(g["f"]):
            G.eta[q], G.rs[q], G.col[q] = eta, rs, col
    t.n_ag = len(angular)
    for gi, g in enumerate(angular):
        G = t.ag[gi]
        G.ej, G.ek, G.cut, G.rc, G.kind = g["key"]
        G.half = 0.5 if G.ej == G.ek else 1.0
        G.n_eta, G.n_ang = len(g["etas"]), len(g["angs"])
        for e, eta in enumerate(g["etas"]):
            G.eta[e] = eta
        for a, (zeta, lam) in enumerate(g["angs"]):
            G.zeta[a], G.lam[a] = zeta, lam
            G.pre[a] = 2.0 ** (1.0 - zeta)
            G.izeta[a] = int(zeta) if (float(zeta).is_integer() and 0 <= zeta <= 16) else -1
        for k in range(_lib.PT_AG_ETA * _lib.PT_AG_ANG):
            G.col[k] = -1
        for (e, a), col in g["cells"].items():
            G.col[e * _lib.PT_AG_ANG + a] = col
    return t


class DeviceTable:
    """A compiled pt_sf_table uploaded to device memory."""

    def __init__(self, rows: Sequence[SFRow], device):
        torch = _torch()
        self.rows = tuple(rows)
        self.n_sf = len(rows)
        self.rc_max = max([r.rc for r in rows], default=0.0)
        self.host = compile_table(rows)
        raw = bytes(self.host)
        self.buf = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(device)
        # kernel variant for this table class (pt_sf_table_kind of the same bytes)
        self.kind = int(_lib.load().pt_sf_table_kind(ctypes.byref(self.host)))

    @property
    def ptr(self) -> int:
        return self.buf.data_ptr()


# ---------------------------------------------------------------------------
# kernel wrappers
# ---------------------------------------------------------------------------

PRECISIONS = {"float64": _lib.PT_F64, "float32": _lib.PT_F32}


def _dtype(precision: int):
    torch = _torch()
    return torch.float64 if precision == _lib.PT_F64 else torch.float32


def acsf_forward(nl: NeighborList, table: DeviceTable, precision: int, n_centers: int,
                 center_begin: int = 0, centers=None, out=None, ld: Optional[int] = None):
    torch = _torch()
    ld = table.n_sf if ld is None else ld
    if out is None:
        out = torch.zeros((n_centers, ld), dtype=_dtype(precision), device=nl.pos4.device)
    box_p, _keep = _box_arg(nl.box_lengths)
    _lib.check(_lib.load().pt_acsf_forward(
        precision | (table.kind << _lib.PT_ACSF_KIND_SHIFT), nl.pos4.data_ptr(), box_p, nl.row_off.data_ptr(), nl.nbr.data_ptr(), nl.max_row,
        n_centers, center_begin, _vp(centers), table.ptr, out.data_ptr(), ld, _lib.stream_ptr()),
        "pt_acsf_forward")
    return out


def acsf_adjoint(nl: NeighborList, table: DeviceTable, precision: int, n_centers: int, w, ldw: int,
                 center_begin: int = 0, centers=None, f_self=None, edge_grad=None, transposed: bool = False):
    """transposed: edge gradients go to the reverse edge's slot (pt_acsf_adjoint_transposed),
    for force_gather_transposed."""
    box_p, _keep = _box_arg(nl.box_lengths)
    if transposed:
        _lib.check(_lib.load().pt_acsf_adjoint_transposed(
            precision | (table.kind << _lib.PT_ACSF_KIND_SHIFT), nl.pos4.data_ptr(), box_p, nl.row_off.data_ptr(),
            nl.nbr.data_ptr(), nl.rev.data_ptr(), nl.max_row, n_centers, center_begin, _vp(centers), table.ptr,
            w.data_ptr(), ldw, _vp(f_self), _vp(edge_grad), _lib.stream_ptr()), "pt_acsf_adjoint_transposed")
        return
    _lib.check(_lib.load().pt_acsf_adjoint(
        precision | (table.kind << _lib.PT_ACSF_KIND_SHIFT), nl.pos4.data_ptr(), box_p, nl.row_off.data_ptr(), nl.nbr.data_ptr(), nl.max_row,
        n_centers, center_begin, _vp(centers), table.ptr, w.data_ptr(), ldw, _vp(f_self),
        _vp(edge_grad), _lib.stream_ptr()), "pt_acsf_adjoint")


def acsf_tangent(nl: NeighborList, table: DeviceTable, precision: int, n_centers: int, delta,
                 center_only: bool, center_begin: int = 0, centers=None, out=None,
                 ld: Optional[int] = None):
    torch = _torch()
    ld = table.n_sf if ld is None else ld
    if out is None:
        out = torch.zeros((n_centers, ld), dtype=_dtype(precision), device=nl.pos4.device)
    box_p, _keep = _box_arg(nl.box_lengths)
    _lib.check(_lib.load().pt_acsf_tangent(
        precision | (table.kind << _lib.PT_ACSF_KIND_SHIFT), nl.pos4.data_ptr(), box_p, nl.row_off.data_ptr(),
        nl.nbr.data_ptr(), nl.max_row,
        n_centers, center_begin, _vp(centers), table.ptr, delta.data_ptr(), int(center_only),
        out.data_ptr(), ld, _lib.stream_ptr()), "pt_acsf_tangent")
    return out


def force_gather(nl: NeighborList, edge_grad, precision: int, out=None):
    torch = _torch()
    if out is None:
        out = torch.empty((nl.n, 3), dtype=torch.float64, device=nl.pos4.device)
    _lib.check(_lib.load().pt_force_gather(nl.n, nl.row_off.data_ptr(), nl.rev.data_ptr(),
                                           edge_grad.data_ptr(), precision, out.data_ptr(),
                                           _lib.stream_ptr()), "pt_force_gather")
    return out


def force_gather_transposed(nl: NeighborList, edge_grad_t, precision: int, f_self, out=None):
    """F = f_self - row sums of the transposed edge buffer (pt_force_gather_transposed)."""
    torch = _torch()
    if out is None:
        out = torch.empty((nl.n, 3), dtype=torch.float64, device=nl.pos4.device)
    _lib.check(_lib.load().pt_force_gather_transposed(nl.n, nl.row_off.data_ptr(), f_self.data_ptr(),
                                                      edge_grad_t.data_ptr(), precision, out.data_ptr(),
                                                      _lib.stream_ptr()), "pt_force_gather_transposed")
    return out


def gather_rows(idx, src, out=None):
    """out[i] = src[idx[i]] (rows of a 2-D FP64 tensor)."""
    torch = _torch()
    n = int(idx.shape[0])
    width = int(src.shape[1]) if src.dim() == 2 else 1
    if out is None:
        out = torch.empty((n, width) if src.dim() == 2 else (n,), dtype=src.dtype, device=src.device)
    if src.dtype == torch.float64:
        fn = _lib.load().pt_gather_rows_f64
    elif src.dtype == torch.int32:
        fn = _lib.load().pt_gather_rows_i32
    else:
        raise TypeError(src.dtype)
    _lib.check(fn(n, width, idx.data_ptr(), src.contiguous().data_ptr(), out.data_ptr(),
                  _lib.stream_ptr()), "gather_rows")
    return out


def scatter_rows(idx, src, out):
    """out[idx[i]] = src[i] (rows of a 2-D FP64 tensor)."""
    n = int(idx.shape[0])
    width = int(src.shape[1]) if src.dim() == 2 else 1
    _lib.check(_lib.load().pt_scatter_rows_f64(n, width, idx.data_ptr(), src.contiguous().data_ptr(),
                                               out.data_ptr(), _lib.stream_ptr()), "scatter_rows")
    return out


class Reducer:
    """Deterministic FP64 sums on the device (pt_reduce_sum_f64)."""

    def __init__(self, n: int, device):
        torch = _torch()
        self.n = n
        self.ws = torch.empty(_lib.load().pt_reduce_workspace_bytes(max(n, 1)), dtype=torch.uint8,
                              device=device)
        self.out = torch.zeros(4, dtype=torch.float64, device=device)

    def sum(self, x, slot: int = 0):
        _lib.check(_lib.load().pt_reduce_sum_f64(int(x.numel()), x.data_ptr(), self.out[slot:].data_ptr(),
                                                 self.ws.data_ptr(), self.ws.numel(), _lib.stream_ptr()),
                   "pt_reduce_sum_f64")
        return self.out[slot]

    def kinetic(self, vel, mass, slot: int = 1):
        _lib.check(_lib.load().pt_kinetic_energy(int(vel.shape[0]), vel.data_ptr(), mass.data_ptr(),
                                                 self.out[slot:].data_ptr(), self.ws.data_ptr(),
                                                 self.ws.numel(), _lib.stream_ptr()), "pt_kinetic_energy")
        return self.out[slot]
