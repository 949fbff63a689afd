import json
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the gfx950 kernels")


@pytest.fixture(scope="session", autouse=True)
def _oracle_built():
    lib = ROOT / "oracle" / "liboracle.so"
    if not lib.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    yield


@pytest.fixture(scope="session")
def goldens():
    return json.loads((GOLDEN / "reference_goldens.json").read_text())


def read_runner_host(path, index=0):
    from pantea_amd.io import read_runner_blocks

    for i, b in enumerate(read_runner_blocks(path)):
        if i == index:
            return b
    raise IndexError(index)


def h2o_fixture_inputs(dtype=np.float32):
    """Structure 0 of h2o.data and the h2o.json potential, as oracle inputs.

    test_nnp.py:17 evaluates the reference in float32: positions, scaler
    parameters and weights are rounded to float32 first.
    """
    from oracle import pyoracle

    b = read_runner_host(GOLDEN / "h2o.data", 0)
    L = np.asarray(b["lattice"], dtype=dtype)
    pos = np.asarray(b["positions"], dtype=dtype)
    box = np.diag(L)
    pos = np.remainder(pos, box).astype(dtype).astype(np.float64)
    types = np.array([2 if e == "O" else 1 for e in b["elements"]], dtype=np.int32)
    s = json.loads((GOLDEN / "h2o.json").read_text())
    tmap = {"H": 1, "O": 2}
    code = pyoracle.CUTOFF_CODES[s["cutoff_type"]]
    tables = {"H": [], "O": []}
    for a in s["symfunction_short"]:
        if a[1] < 3:
            c, k, ej, eta, rc, rs = a
            tables[c].append([k, tmap[ej], 0, eta, rs, 0, 0, rc, code, 0])
        else:
            c, k, ej, eta, rc, rs, ek, lam, zeta = a
            tables[c].append([k, tmap[ej], tmap[ek], eta, rs, zeta, lam, rc, code, 0])
    scal, layers = {}, {}
    acts = s["global_activation_short"]
    for e, z in (("H", "001"), ("O", "008")):
        p = json.loads((GOLDEN / f"scaling.{z}.json").read_text())
        scal[e] = {k: np.asarray(v, dtype=dtype).astype(np.float64) for k, v in p.items()}
        w = np.load(GOLDEN / f"weights.{z}.npz")
        layers[e] = [(w[f"layers_{2*i}/kernel"].astype(np.float64), w[f"layers_{2*i}/bias"].astype(np.float64), acts[i])
                     for i in range(len(acts))]
    nnp = pyoracle.OracleNNP(["H", "O"], tmap, tables, s["scale_type"], scal, layers)
    return pos, types, L.astype(np.float64), nnp
