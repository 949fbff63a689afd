"""Per-element atomic MLP (models/nn/model.py:16-81) on gfx950 matrix cores.

Parameters keep the reference's flax layout: ``{"layers_0": {"kernel":
(in, out), "bias": (out,)}, "layers_2": ..., ...}`` with y = x @ W + b
(model.py:28-38) -- the odd layer indices are the activations.  Evaluation
runs pt_mlp_energy_grad (FP64 MFMA, or BF16 / split-BF16 MFMA with FP32
accumulation).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from . import _lib

ModelParams = Dict[str, Dict[str, np.ndarray]]

MLP_MODES = {"float64": _lib.PT_MLP_F64, "bf16": _lib.PT_MLP_BF16, "bf16x3": _lib.PT_MLP_BF16X3}


class UniformInitializer:
    """Uniform[w_min, w_max) kernel initialiser (models/nn/initializer.py:10-20)."""

    def __init__(self, weights_range: Tuple[float, float]):
        self.weights_range = tuple(weights_range)

    def __call__(self, rng: np.random.Generator, shape) -> np.ndarray:
        lo, hi = self.weights_range
        return rng.uniform(lo, hi, size=shape)


class NormalInitializer:
    """N(0, scale^2 / fan_in) kernel initialiser (scale 1: variance 1/fan_in)."""

    def __init__(self, scale: float = 1.0):
        self.scale = float(scale)

    def __call__(self, rng: np.random.Generator, shape) -> np.ndarray:
        return rng.normal(0.0, self.scale / np.sqrt(shape[0]), size=shape)


class LecunNormal:
    """Truncated normal with variance 1/fan_in (flax lecun_normal semantics)."""

    def __call__(self, rng: np.random.Generator, shape) -> np.ndarray:
        std = 1.0 / np.sqrt(shape[0]) / 0.87962566103423978
        x = rng.normal(size=shape)
        bad = np.abs(x) > 2.0
        while np.any(bad):
            x[bad] = rng.normal(size=int(bad.sum()))
            bad = np.abs(x) > 2.0
        return x * std


@dataclass
class NeuralNetworkModel:
    hidden_layers: Tuple[Tuple[int, str], ...]
    output_layer: Tuple[int, str] = (1, "identity")
    params_dtype: str = "float64"
    kernel_initializer: object = field(default_factory=LecunNormal)

    def __post_init__(self) -> None:
        self.hidden_layers = tuple((int(n), str(a)) for n, a in self.hidden_layers)
        self.output_layer = (int(self.output_layer[0]), str(self.output_layer[1]))
        for _, a in self.hidden_layers + (self.output_layer,):
            if a not in _lib.ACTIVATION_CODES:
                raise KeyError(f"unknown activation '{a}'")

    def init(self, seed: int, in_dim: int) -> ModelParams:
        """Random kernels (initialiser), zero biases (model.py:28-38)."""
        rng = np.random.default_rng(seed)
        params: ModelParams = {}
        fan_in = in_dim
        widths = [n for n, _ in self.hidden_layers] + [self.output_layer[0]]
        for i, w in enumerate(widths):
            params[f"layers_{2 * i}"] = {
                "kernel": np.asarray(self.kernel_initializer(rng, (fan_in, w)), dtype=np.float64),
                "bias": np.zeros(w, dtype=np.float64),
            }
            fan_in = w
        return params

    def desc(self, in_dim: int, scale_add: float = 0.0) -> _lib.MlpDesc:
        d = _lib.MlpDesc()
        layers = self.hidden_layers + (self.output_layer,)
        if self.output_layer[0] != 1:
            raise ValueError("the output layer must have width 1")
        d.n_layers = len(layers)
        d.in_dim = in_dim
        for i, (w, a) in enumerate(layers):
            d.width[i] = w
            d.act[i] = _lib.ACTIVATION_CODES[a]
        d.scale_add = scale_add
        return d

    @staticmethod
    def flatten(params: ModelParams) -> np.ndarray:
        keys = sorted(params, key=lambda k: int(k.split("_")[1]))
        parts = []
        for k in keys:
            parts.append(np.asarray(params[k]["kernel"], dtype=np.float64).ravel())
            parts.append(np.asarray(params[k]["bias"], dtype=np.float64).ravel())
        return np.concatenate(parts)

    def save(self, filename, params: ModelParams) -> None:
        """Weights as .npz (``layers_{i}/kernel``, ``layers_{i}/bias``)."""
        flat = {f"{k}/{p}": np.asarray(v) for k, d in params.items() for p, v in d.items()}
        np.savez(str(filename), **flat)

    def load(self, filename) -> ModelParams:
        """Load .npz weights (the reference pickles them, model.py:68-81; pickles are not read here)."""
        with np.load(str(filename), allow_pickle=False) as z:
            out: ModelParams = {}
            for key in z.files:
                layer, p = key.split("/")
                out.setdefault(layer, {})[p] = np.asarray(z[key], dtype=np.float64)
        return out

    def __repr__(self) -> str:
        return f"NeuralNetworkModel(hidden_layers={self.hidden_layers}, dtype={self.params_dtype})"


class DeviceMLP:
    """Packed device copy of one element's MLP (pt_mlp_pack) plus its scaler."""

    def __init__(self, model: NeuralNetworkModel, params: ModelParams, in_dim: int,
                 scale_sub_mul: np.ndarray, scale_add: float, device):
        import torch

        self.model = model
        self.in_dim = in_dim
        self.desc = model.desc(in_dim, scale_add)
        lib = _lib.load()
        flat = torch.as_tensor(NeuralNetworkModel.flatten(params), dtype=torch.float64, device=device)
        nbytes = lib.pt_mlp_packed_bytes(ctypes.byref(self.desc))
        if nbytes == 0:
            raise _lib.HipLibraryError(f"unsupported MLP shape: {_lib.last_error()}")
        self.packed = torch.zeros(nbytes, dtype=torch.uint8, device=device)
        _lib.check(lib.pt_mlp_pack(ctypes.byref(self.desc), flat.data_ptr(), self.packed.data_ptr(),
                                   _lib.stream_ptr()), "pt_mlp_pack")
        self.flat = flat  # FP64 master parameters (W then b per layer)
        self.scale = torch.as_tensor(scale_sub_mul, dtype=torch.float64, device=device).contiguous()
        self.n_params = int(lib.pt_mlp_param_count(ctypes.byref(self.desc)))

    def repack(self) -> None:
        """Rebuild the packed copies after ``flat`` changed (e.g. an optimiser step)."""
        _lib.check(_lib.load().pt_mlp_pack(ctypes.byref(self.desc), self.flat.data_ptr(), self.packed.data_ptr(),
                                           _lib.stream_ptr()), "pt_mlp_pack")

    def params_dict(self) -> ModelParams:
        """The master parameters back in the layers_{2i}/{kernel,bias} layout (host)."""
        flat = self.flat.detach().cpu().numpy()
        out: ModelParams = {}
        fan_in = self.in_dim
        o = 0
        widths = [w for w, _ in self.model.hidden_layers] + [self.model.output_layer[0]]
        for i, w in enumerate(widths):
            W = flat[o:o + fan_in * w].reshape(fan_in, w)
            o += fan_in * w
            b = flat[o:o + w]
            o += w
            out[f"layers_{2 * i}"] = {"kernel": W.copy(), "bias": b.copy()}
            fan_in = w
        return out

    def train_grad(self, mode: int, n_atoms: int, G, T, ldg: int, g_precision: int, gy, tseed: float, grad) -> None:
        _lib.check(_lib.load().pt_mlp_train_grad(
            mode, n_atoms, G.data_ptr(), T.data_ptr(), ldg, g_precision, self.scale.data_ptr(),
            ctypes.byref(self.desc), self.packed.data_ptr(), gy.data_ptr(), float(tseed), grad.data_ptr(),
            _lib.stream_ptr()), "pt_mlp_train_grad")

    def energy_grad(self, mode: int, n_atoms: int, G, ldg: int, g_precision: int, e_atom, dEdG) -> None:
        _lib.check(_lib.load().pt_mlp_energy_grad(
            mode, n_atoms, G.data_ptr(), ldg, g_precision, self.scale.data_ptr(), ctypes.byref(self.desc),
            self.packed.data_ptr(), e_atom.data_ptr(), None if dEdG is None else dEdG.data_ptr(),
            _lib.stream_ptr()), "pt_mlp_energy_grad")
