"""Global extended Kalman filter trainer on the GPU.

Mirrors ``KalmanFilter`` (potentials/nnp/kalman_filter.py:42-335): one
state vector W of every element's MLP weights, a dense error covariance P
(n_w x n_w, FP64, resident in HBM), and per structure either an energy
update (one observation, (E* - E) / N) or a force update (3N observations,
F* - F with the reference's self-term forces, force.py:28-43).  The choice
between them, the shuffling and the schedules of eta, q and lambda follow
the reference step for step, with the same random generators (Python's
``random`` for the shuffle, numpy's legacy generator for the energy/force
choice, both seeded with ``random_seed``).

Per structure the descriptors G and the centre-only tangents T_a = dG_i/dr_ia
(a = x, y, z) do not depend on the weights and are built once.  An update
then runs

    energy:  pt_mlp_energy_grad, pt_reduce_sum_f64, pt_kalman_residual,
             H = (1/N) sum_i dE_i/dW           pt_mlp_train_grad
    force:   pt_mlp_energy_grad, pt_acsf_adjoint (self term),
             pt_kalman_residual, H = dF/dW      pt_mlp_jacobian
    both:    pt_kalman_update (rocBLAS GEMMs + rocSOLVER Cholesky), pt_mlp_pack

with nothing read back to the host inside an epoch.
"""

from __future__ import annotations

import ctypes
import math
import random
from collections import defaultdict
from dataclasses import dataclass, field
from typing import Any, Callable, Dict, List, Optional, Tuple

import numpy as np

from . import _lib
from .engine import NeighborList, Reducer, acsf_adjoint, acsf_forward, acsf_tangent, gather_rows
from .potential import NNPSettings, NeuralNetworkPotential


@dataclass
class TrainingParams:
    """trainer.py:29-44."""

    force_weight: float
    energy_fraction: float
    force_fraction: float
    epochs: int

    @classmethod
    def from_runner(cls, settings: NNPSettings) -> "TrainingParams":
        return cls(force_weight=settings.force_weight, energy_fraction=settings.short_energy_fraction,
                   force_fraction=settings.short_force_fraction, epochs=settings.epochs)


@dataclass
class _Item:
    """One labelled structure on the device, internal atom order."""

    n: int
    nl: NeighborList
    blocks: List[Tuple[str, int, int]]
    G: Any          # (n, ld) FP64 descriptors
    T: Any          # (3, n, ld) FP64 centre-only tangents
    F_ref: Any      # (n, 3) FP64
    E_ref: float


@dataclass
class KalmanFilter:
    """kalman_filter.py:42-110 (state and settings) and :112-233 (fit, update)."""

    kalman_type: int
    epsilon: float
    q: float
    q_tau: float
    q_min: float
    eta: float
    eta_tau: float
    eta_max: float
    lambda0: float
    neu: float
    gamma: float
    num_states: int
    W: Any = field(repr=False)
    P: Any = field(repr=False)
    potential: NeuralNetworkPotential = field(repr=False)
    random_seed: int = 2023
    offsets: Dict[str, int] = field(default_factory=dict, repr=False)

    # ---- construction ----------------------------------------------------------------
    @classmethod
    def from_runner(cls, potential: NeuralNetworkPotential, filename: str = "input.nn") -> "KalmanFilter":
        settings = NNPSettings.from_file(potential.directory / filename)
        return cls.from_settings(potential, settings)

    @classmethod
    def from_settings(cls, potential: NeuralNetworkPotential, settings: NNPSettings) -> "KalmanFilter":
        import torch

        dev = _lib.require_device()
        offsets, parts, o = {}, [], 0
        for e in potential.elements:
            flat = np.concatenate([np.asarray(p, dtype=np.float64).ravel()
                                   for k in sorted(potential.models_params[e], key=lambda k: int(k.split("_")[1]))
                                   for p in (potential.models_params[e][k]["kernel"],
                                             potential.models_params[e][k]["bias"])])
            offsets[e] = o
            parts.append(flat)
            o += flat.size
        n_w = o
        if n_w * n_w * 8 > (64 << 30):
            raise MemoryError(f"the covariance of {n_w} weights needs {n_w * n_w * 8 / 2**30:.0f} GiB")
        W = torch.as_tensor(np.concatenate(parts), dtype=torch.float64, device=dev)
        P = torch.zeros((n_w, n_w), dtype=torch.float64, device=dev)
        P.diagonal().fill_(1.0 / settings.kalman_epsilon)
        return cls(kalman_type=int(settings.kalman_type), epsilon=settings.kalman_epsilon, q=settings.kalman_q0,
                   q_tau=settings.kalman_qtau, q_min=settings.kalman_qmin, eta=settings.kalman_eta,
                   # kalman_filter.py:98 initialises eta_tau from kalman_eta, not kalman_etatau
                   eta_tau=settings.kalman_eta, eta_max=settings.kalman_etamax,
                   lambda0=settings.kalman_lambda_short, neu=settings.kalman_neu_short, gamma=1.0,
                   num_states=n_w, W=W, P=P, potential=potential, random_seed=settings.random_seed,
                   offsets=offsets)

    # ---- device state --------------------------------------------------------------------
    def _bind(self) -> None:
        """Device MLPs whose FP64 master parameters are views into W."""
        pot = self.potential
        pot._check_scaler_params_exist()
        pot._mlp_key = None
        self.mlps = pot._device_mlps(self.dev)
        for e, m in self.mlps.items():
            o = self.offsets[e]
            if o + m.n_params > self.num_states:
                raise RuntimeError("parameter layout does not match the state vector")
            m.flat = self.W[o:o + m.n_params]
            m.repack()

    def _prepare(self, dataset) -> List[_Item]:
        import torch

        pot = self.potential
        ld = max(ap.model_input_size for ap in pot.atomic_potentials.values())
        self.ld = ld
        items = []
        for index in range(len(dataset)):
            s = dataset[index]
            emap = s.element_map
            missing = [e for e in emap.unique_elements if e not in pot.atomic_potentials]
            if missing:
                raise ValueError(f"structure {index} holds elements without a model: {missing}")
            nl = NeighborList(pot.r_cutoff, emap.n_types).build(s.positions, s.atom_types, s.box_lengths,
                                                                s.type_counts(), with_rev=True)
            n = s.natoms
            G = torch.zeros((n, ld), dtype=torch.float64, device=self.dev)
            T = torch.zeros((3, n, ld), dtype=torch.float64, device=self.dev)
            blocks = []
            for e in pot.elements:
                t = emap.element_to_atom_type.get(e)
                if t is None:
                    continue
                b, cnt = nl.type_block(t)
                if cnt == 0:
                    continue
                tab = pot.atomic_potentials[e].descriptor.device_table(emap, self.dev)
                acsf_forward(nl, tab, _lib.PT_F64, cnt, center_begin=b, out=G[b:], ld=ld)
                for a in range(3):
                    delta = torch.zeros((n, 3), dtype=torch.float64, device=self.dev)
                    delta[:, a] = 1.0
                    acsf_tangent(nl, tab, _lib.PT_F64, cnt, delta, center_only=True, center_begin=b,
                                 out=T[a, b:], ld=ld)
                blocks.append((e, b, cnt, tab))
            F_ref = gather_rows(nl.perm, s.forces.to(torch.float64).contiguous())
            nl.release_workspace()
            items.append(_Item(n, nl, blocks, G, T, F_ref, float(s.total_energy)))
        return items

    def _buffers(self, items: List[_Item]) -> None:
        import torch

        dev, n_w = self.dev, self.num_states
        nmax = max(it.n for it in items)
        self.e_atom = torch.zeros(nmax, dtype=torch.float64, device=dev)
        self.dEdG = torch.zeros((nmax, self.ld), dtype=torch.float64, device=dev)
        self.f_self = torch.zeros((nmax, 3), dtype=torch.float64, device=dev)
        self.gy = torch.zeros(nmax, dtype=torch.float64, device=dev)
        self.xi = torch.zeros(3 * nmax, dtype=torch.float64, device=dev)
        self.H = torch.zeros((3 * nmax, n_w), dtype=torch.float64, device=dev)
        self.loss = torch.zeros(2, dtype=torch.float64, device=dev)
        self.info = torch.zeros(1, dtype=torch.int32, device=dev)
        self.info_sum = torch.zeros(1, dtype=torch.int64, device=dev)
        lib = _lib.load()
        self.ws = torch.empty(lib.pt_kalman_workspace_bytes(n_w, 3 * nmax), dtype=torch.uint8, device=dev)
        self.reducer = Reducer(nmax, dev)

    # ---- observations ----------------------------------------------------------------------
    def _energy(self, it: _Item):
        for e, b, cnt, _ in it.blocks:
            self.mlps[e].energy_grad(_lib.PT_MLP_F64, cnt, it.G[b:], self.ld, _lib.PT_F64, self.e_atom[b:],
                                     self.dEdG[b:])
        return self.reducer.sum(self.e_atom[:it.n])

    def calculate_loss_energy(self, it: _Item):
        """kalman_filter.py:235-249: xi = (E* - E) / N, H = -dxi/dW = (1/N) sum_i dE_i/dW."""
        lib = _lib.load()
        E = self._energy(it)
        _lib.check(lib.pt_kalman_residual(1, E.data_ptr(), None, it.E_ref, 1.0 / it.n, self.xi.data_ptr(),
                                          self.loss[0:].data_ptr(), _lib.stream_ptr()), "pt_kalman_residual")
        H = self.H[:1]
        H.zero_()
        self.gy[:it.n].fill_(1.0 / it.n)
        for e, b, cnt, _ in it.blocks:
            m = self.mlps[e]
            m.train_grad(_lib.PT_MLP_F64, cnt, it.G[b:], it.G[b:], self.ld, _lib.PT_F64, self.gy[b:], 0.0,
                         H[0, self.offsets[e]:])
        return self.xi[:1], H

    def calculate_loss_force(self, it: _Item):
        """kalman_filter.py:251-273: xi = F* - F (self-term forces), H = dF/dW (3N rows)."""
        lib = _lib.load()
        self._energy(it)
        f = self.f_self[:it.n]
        f.zero_()
        for e, b, cnt, tab in it.blocks:
            acsf_adjoint(it.nl, tab, _lib.PT_F64, cnt, self.dEdG[b:], self.ld, center_begin=b, f_self=f[b:])
        m3 = 3 * it.n
        _lib.check(lib.pt_kalman_residual(m3, f.data_ptr(), it.F_ref.data_ptr(), 0.0, 1.0, self.xi.data_ptr(),
                                          self.loss[1:].data_ptr(), _lib.stream_ptr()), "pt_kalman_residual")
        H = self.H[:m3]
        H.zero_()
        n_w = self.num_states
        for e, b, cnt, _ in it.blocks:
            m = self.mlps[e]
            J = H[3 * b:]
            _lib.check(lib.pt_mlp_jacobian(cnt, 3, it.G[b:].data_ptr(), it.T[:, b:].data_ptr(), it.n * self.ld,
                                           self.ld, m.scale.data_ptr(), ctypes.byref(m.desc), m.packed.data_ptr(),
                                           -1.0, J.data_ptr() + 8 * self.offsets[e], n_w, _lib.stream_ptr()),
                       "pt_mlp_jacobian")
        return self.xi[:m3], H

    # ---- the filter ----------------------------------------------------------------------------
    def update_state_vector(self, Xi, H) -> None:
        """kalman_filter.py:196-233 (schedules on the host, linear algebra on the device)."""
        if self.kalman_type == 0 and self.eta < self.eta_max:
            self.eta *= math.exp(self.eta_tau)
        if self.kalman_type == 0:
            r, c = 1.0 / self.eta, 1.0
        elif self.kalman_type == 1:
            r, c = self.lambda0, 1.0 / self.lambda0
        else:
            raise ValueError(f"unknown kalman_type {self.kalman_type}")
        m = int(Xi.shape[0])
        lib = _lib.load()
        _lib.check(lib.pt_kalman_update(self.num_states, m, self.P.data_ptr(), H.data_ptr(), self.num_states,
                                        Xi.data_ptr(), self.W.data_ptr(), r, c, self.q, self.ws.data_ptr(),
                                        self.ws.numel(), self.info.data_ptr(), _lib.stream_ptr()),
                   "pt_kalman_update")
        self.info_sum.add_(self.info)
        if self.q > self.q_min:
            self.q *= math.exp(-self.q_tau)
        if self.kalman_type == 1:
            self.lambda0 = self.neu * self.lambda0 + 1.0 - self.neu
            self.gamma = 1.0 / (1.0 + self.lambda0 / self.gamma)
        for e, mlp in self.mlps.items():
            mlp.repack()

    def fit(self, dataset, training_params, log: Optional[Callable[[str], None]] = None):
        """kalman_filter.py:112-194.  Returns the history (epoch, loss, loss_energy, loss_force)."""
        self.dev = _lib.require_device()
        self._bind()
        items = self._prepare(dataset)
        if not items:
            raise ValueError("empty dataset")
        self._buffers(items)
        shuffle_rng = random.Random(self.random_seed)
        choice_rng = np.random.RandomState(self.random_seed)
        indices = list(range(len(items)))
        history: Dict[str, list] = defaultdict(list)
        for epoch in range(training_params.epochs):
            shuffle_rng.shuffle(indices)
            self.loss.zero_()
            n_e = n_f = 0
            for index in indices:
                it = items[index]
                if choice_rng.rand() < training_params.force_fraction:
                    Xi, H = self.calculate_loss_force(it)
                    n_f += 1
                else:
                    Xi, H = self.calculate_loss_energy(it)
                    n_e += 1
                self.update_state_vector(Xi, H)
            lo = self.loss.cpu().numpy()
            if int(self.info_sum.item()) != 0:
                raise _lib.HipLibraryError("Kalman update: the innovation covariance lost positive definiteness")
            le = lo[0] / (n_e if n_e > 0 else 1)
            lf = lo[1] / (n_f if n_f > 0 else 1)
            self.sync_params()
            history["epoch"].append(epoch + 1)
            history["loss"].append(le + lf)
            history["loss_energy"].append(le)
            history["loss_force"].append(lf)
            if log:
                log(f"epoch {epoch + 1}: loss {le + lf:.7f}, loss_energy {le:.7f}, loss_force {lf:.7f}, "
                    f"energy_update_ratio {n_e / max(n_e + n_f, 1):.3f}")
        return history

    def sync_params(self) -> None:
        """Write W back into the potential's per-element parameter dicts (kalman_filter.py:165-167)."""
        w = self.W.detach().cpu().numpy()
        params = {}
        for e, m in self.mlps.items():
            o = self.offsets[e]
            flat, out, fan_in = w[o:o + m.n_params], {}, m.in_dim
            p = 0
            widths = [wd for wd, _ in m.model.hidden_layers] + [m.model.output_layer[0]]
            for i, wd in enumerate(widths):
                out[f"layers_{2 * i}"] = {"kernel": flat[p:p + fan_in * wd].reshape(fan_in, wd).copy(),
                                          "bias": flat[p + fan_in * wd:p + fan_in * wd + wd].copy()}
                p += fan_in * wd + wd
                fan_in = wd
            params[e] = out
        # keep the bound device MLPs (views into W) valid for the next epoch
        self.potential.models_params = params
        key = (id(self.potential.models_params), str(self.dev),
               tuple(id(v) for v in self.potential.scalers_params.values()))
        self.potential._mlp_key = key


@dataclass
class NeuralNetworkPotentialTrainer:
    """trainer.py:47-147 with the Kalman filter as the updater."""

    params: TrainingParams
    updater: Any
    potential: NeuralNetworkPotential = field(repr=False)

    @classmethod
    def from_runner(cls, potential: NeuralNetworkPotential, filename: str = "input.nn"):
        settings = NNPSettings.from_file(potential.directory / filename)
        return cls.from_settings(potential, settings)

    @classmethod
    def from_settings(cls, potential: NeuralNetworkPotential, settings: NNPSettings):
        if settings.updater_type != "kalman_filter":
            raise NotImplementedError(
                f"updater_type {settings.updater_type!r}: the device trainer here is the Kalman filter; "
                "Adam training of a potential is pantea_amd.student.StudentTrainer")
        return cls(TrainingParams.from_runner(settings), KalmanFilter.from_settings(potential, settings), potential)

    def fit_scaler(self, dataset) -> None:
        self.potential.fit_scaler([dataset[i] for i in range(len(dataset))])

    def fit_model(self, dataset, log=None):
        return self.updater.fit(dataset, self.params, log=log)

    def fit(self, dataset, log=None):
        self.fit_scaler(dataset)
        return self.fit_model(dataset, log=log)

    def save(self, directory=None) -> None:
        self.potential.save(directory)
