"""Convert the reference's own test fixtures into plain data files.

Run once, in the build container, from the repository root:

    python tests/golden/convert_reference_fixtures.py

Inputs (reference tree, read-only):
  tests/h2o.data, tests/h2o.json, tests/scaling.001.json,
  tests/scaling.008.json                    -> copied verbatim (text data)
  tests/weights.001.pkl, tests/weights.008.pkl -> weights.00Z.npz

The weight pickles hold float32 numpy payloads wrapped in
``jax._src.array._reconstruct_array``.  ``torch.load(weights_only=True)``
refuses them (the jax global is not allow-listed), and no unpickler is used
here.  Instead the opcode stream is walked statically with
``pickletools.genops``: that parser only decodes opcodes and their literal
arguments; it constructs no object, imports no module and resolves no
global.  The walk collects, in stream order, the layer names
(``layers_0`` ...), parameter names (``kernel``/``bias``), the shape tuples
and the raw little-endian float32 byte strings.

Nothing here is imported by the package or by the tests; the tests read the
files this script writes.
"""

from __future__ import annotations

import json
import pickletools
import shutil
import sys
from pathlib import Path

import numpy as np

REF = Path("/root/reference/tests")
OUT = Path(__file__).resolve().parent


def decode_weights(path: Path) -> dict:
    data = path.read_bytes()
    layer = None
    param = None
    arrays = {}
    memo: dict = {}
    last = None  # literal pushed by the previous opcode (for MEMOIZE/BINPUT)

    def see(name):
        nonlocal layer, param
        if isinstance(name, str):
            if name.startswith("layers_"):
                layer = name
            elif name in ("kernel", "bias"):
                param = name

    for op, arg, _pos in pickletools.genops(data):
        if op.name in ("SHORT_BINUNICODE", "BINUNICODE", "UNICODE"):
            see(arg)
            last = arg
            continue
        if op.name == "MEMOIZE":
            memo[len(memo)] = last
            continue
        if op.name in ("BINPUT", "LONG_BINPUT", "PUT"):
            memo[arg] = last
            continue
        if op.name in ("BINGET", "LONG_BINGET", "GET"):
            last = memo.get(arg)
            see(last)
            continue
        last = None
        if op.name in ("SHORT_BINBYTES", "BINBYTES", "BINBYTES8"):
            if len(arg) <= 1:
                continue  # numpy's b'b' placeholder of ndarray._reconstruct
            if layer is None or param is None:
                raise ValueError(f"payload before names in {path}")
            vals = np.frombuffer(arg, dtype="<f4").copy()
            arrays[f"{layer}/{param}"] = vals
            param = None
    # Shapes: flax Dense kernel is (fan_in, fan_out); bias is (fan_out,).
    out = {}
    for key, vals in arrays.items():
        layer, param = key.split("/")
        if param == "bias":
            out[key] = vals
    for key, vals in arrays.items():
        layer, param = key.split("/")
        if param == "kernel":
            fan_out = out[f"{layer}/bias"].size
            out[key] = vals.reshape(vals.size // fan_out, fan_out)
    return out


def main() -> int:
    if not REF.exists():
        print("reference tree not present; nothing to convert", file=sys.stderr)
        return 1
    for name in ("h2o.data", "h2o.json", "scaling.001.json", "scaling.008.json"):
        shutil.copyfile(REF / name, OUT / name)
    for z in ("001", "008"):
        arrs = decode_weights(REF / f"weights.{z}.pkl")
        np.savez(OUT / f"weights.{z}.npz", **arrs)
        print(z, {k: v.shape for k, v in arrs.items()})
    return 0


if __name__ == "__main__":
    sys.exit(main())
