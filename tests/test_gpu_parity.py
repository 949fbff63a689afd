"""GPU parity: the gfx950 kernels (through the C ABI) against the FP64 oracle.

Bars: neighbour lists bit-exact (same directed pairs); FP64 descriptors,
energies and forces within 1e-10 relative (scaled by the largest magnitude
of the compared array); FP32/BF16 paths within the error budgets stated in
each test.
"""

import json
import sys

import numpy as np
import pytest

from conftest import GOLDEN, h2o_fixture_inputs

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from pantea_amd import _lib

    _lib.load()
    return torch.device("cuda")


def _rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _structure(data):
    from pantea_amd.atoms import Structure

    return Structure.from_dict(data)


def _water(name):
    from pantea_amd import water

    return _structure(water.config(name))


# ---------------------------------------------------------------------------
# (a) neighbour list
# ---------------------------------------------------------------------------

def _rc_boundary_pairs(rc):
    """72 isolated pairs whose separations are rc, rc +- a few ulps (axis, face
    and body diagonals), in a periodic box of 60."""
    pts = []
    for p in range(72):
        base = np.array([10.0 * (p % 6), 10.0 * ((p // 6) % 6), 10.0 * (p // 36)])
        k, kind = p % 9 - 4, (p // 9) % 3
        d = np.array([rc, 0.0, 0.0]) if kind == 0 else (
            np.array([rc, rc, 0.0]) / np.sqrt(2.0) if kind == 1 else np.full(3, rc / np.sqrt(3.0)))
        q = base + d
        for _ in range(abs(k)):
            q[0] = np.nextafter(q[0], np.inf if k > 0 else -np.inf)
        pts += [base, q]
    pos = np.array(pts)
    dd = pos[1::2] - pos[0::2]
    s2 = dd[:, 0] * dd[:, 0]
    s2 = s2 + dd[:, 1] * dd[:, 1]
    s2 = s2 + dd[:, 2] * dd[:, 2]
    r = np.sqrt(s2)
    # the set must hold separations exactly at rc and one ulp either side
    assert np.any(r == rc) and np.any((r > rc) & (r < rc * (1 + 1e-15))) and np.any((r < rc) & (r > rc * (1 - 1e-15)))
    return {"positions": pos, "elements": ["H"] * len(pos), "lattice": np.diag([60.0] * 3)}


@pytest.mark.parametrize("case", ["cfg0_rc12", "box7_rc6", "box10_rc3", "cfg1_rc12", "cfg1_rc8", "cluster_nopbc",
                                  "rc_boundary_ulps", "long_rows", "cluster_far_from_origin",
                                  "clusters_many_candidates", "dense_many_candidates", "sparse_box_dense_cluster"])
def test_neighbor_list_bit_exact(dev, case):
    from oracle import pyoracle
    from pantea_amd import water
    from pantea_amd.engine import NeighborList

    rng = np.random.default_rng(7)
    if case == "cfg0_rc12":
        d, rc = water.config("cfg0"), 12.0
    elif case == "cfg1_rc12":
        d, rc = water.config("cfg1"), 12.0
    elif case == "cfg1_rc8":
        # small cells (~9 atoms, ~250 candidates): the small workgroup configuration
        d, rc = water.config("cfg1"), 8.0
    elif case == "sparse_box_dense_cluster":
        # a sparse box takes the small configuration; the cluster's cells have more candidates
        # than it stages in LDS and read them from HBM
        pos = np.concatenate([rng.random((2000, 3)) * 60.0, 30.0 + rng.uniform(-1.5, 1.5, size=(700, 3))])
        d = {"positions": pos, "elements": ["O", "H"] * 1350, "lattice": np.diag([60.0] * 3)}
        rc = 4.0
    elif case == "box7_rc6":
        d = {"positions": rng.random((40, 3)) * 7.0, "elements": ["O", "H"] * 20, "lattice": np.diag([7.0] * 3)}
        rc = 6.0
    elif case == "box10_rc3":
        d = {"positions": rng.random((300, 3)) * 10.0, "elements": ["H"] * 300, "lattice": np.diag([10.0] * 3)}
        rc = 3.0
    elif case == "rc_boundary_ulps":
        rc = 3.7
        d = _rc_boundary_pairs(rc)
    elif case == "long_rows":
        # ~480 neighbours per atom: longer than the staged row width, so the
        # list takes the second candidate pass
        d = {"positions": rng.random((1200, 3)) * 12.0, "elements": ["O", "H"] * 600, "lattice": np.diag([12.0] * 3)}
        rc = 5.5
    elif case == "clusters_many_candidates":
        # 27 cells of width 10 > rc, each with a tight cluster of 60 atoms:
        # 1620 candidates per cell (more than one workgroup stages in LDS)
        # and rows of 59
        centres = (np.stack(np.meshgrid(*[np.arange(3)] * 3, indexing="ij"), -1).reshape(-1, 1, 3) + 0.5) * 10.0
        pos = (centres + rng.uniform(-0.2, 0.2, size=(27, 60, 3))).reshape(-1, 3)
        d = {"positions": pos, "elements": ["O", "H"] * 810, "lattice": np.diag([30.0] * 3)}
        rc = 9.0
    elif case == "dense_many_candidates":
        # 4000 candidates per cell and rows of ~1600
        d = {"positions": rng.random((4000, 3)) * 12.0, "elements": ["H"] * 4000, "lattice": np.diag([12.0] * 3)}
        rc = 5.5
    elif case == "cluster_far_from_origin":
        # coordinates near 5000: the FP32 candidate screen is coarse (ulp ~5e-4)
        # and many candidates fall to the FP64 test
        d = {"positions": 5000.0 + rng.normal(size=(1500, 3)) * 6.0, "elements": ["O", "H", "H"] * 500,
             "lattice": []}
        rc = 3.5
    else:
        d = {"positions": rng.normal(size=(150, 3)) * 4.0, "elements": ["O", "H", "H"] * 50, "lattice": []}
        rc = 3.5
    _log(f"neighbour case {case}")
    s = _structure(d)
    nl = NeighborList(rc, s.element_map.n_types).build(s.positions, s.atom_types, s.box_lengths, s.type_counts())
    torch.cuda.synchronize()
    counts, idx = nl.to_csr_input_order()
    pos = s.positions.cpu().numpy()
    oc, oi = pyoracle.neighbors(pos, s.box_lengths, rc)
    assert np.array_equal(counts, oc)
    assert np.array_equal(idx, oi)
    # reverse edges
    rev = nl.rev[: nl.n_pairs].cpu().numpy()
    nbr = nl.nbr[: nl.n_pairs].cpu().numpy()
    off = nl.row_off.cpu().numpy()
    assert np.all(rev >= 0)
    assert np.array_equal(rev[rev], np.arange(nl.n_pairs))
    rows_of = np.repeat(np.arange(nl.n), np.diff(off))
    assert np.array_equal(nbr[rev], rows_of)
    # internal order groups types
    t_int = s.atom_types_host[nl.perm_host]
    assert np.all(np.diff(t_int) >= 0)
    assert nl.max_row == int(np.max(counts))
    if case in ("long_rows", "dense_many_candidates"):
        assert nl.max_row > 256
    if case == "clusters_many_candidates":
        assert nl.max_row == 59
    if case == "rc_boundary_ulps":
        assert 0 < nl.n_pairs < 144


# ---------------------------------------------------------------------------
# (b) descriptors
# ---------------------------------------------------------------------------

def test_acsf_goldens_through_api(dev, goldens):
    from pantea_amd.descriptors import ACSF, G2, G3, CutoffFunction, NeighborElements

    g = goldens["acsf_without_pbc"]
    cfn = CutoffFunction.from_type("tanhu", 3.0)
    acsf = ACSF("Ne", tuple((G2(cfn, r_shift=rs, eta=eta), NeighborElements("Ne")) for eta, rs in g["g2"]), ())
    s = _structure({"positions": g["positions"], "elements": g["elements"]})
    G = acsf(s).cpu().numpy()
    assert G.shape == (2, 5)
    assert np.max(np.abs(G - np.tile(g["expected_row"], (2, 1)))) < 1e-10

    g = goldens["acsf_with_pbc"]
    cfn = CutoffFunction.from_type("tanhu", g["cutoff"][1])
    acsf = ACSF("O", ((G2(cfn, r_shift=0.0, eta=0.001), NeighborElements("H")),),
                ((G3(cfn, eta=0.07, zeta=1.0, lambda0=1.0, r_shift=0.0), NeighborElements("H", "H")),))
    s = _structure({"positions": g["positions"], "elements": g["elements"], "lattice": g["lattice"]})
    assert tuple(acsf(s).shape) == (4, 2)
    row = acsf(s, atom_index=0).cpu().numpy()[0]
    assert np.max(np.abs(row - np.array(g["expected_row"]))) < 1e-10
    with pytest.raises(ValueError):
        acsf(s, atom_index=1)  # an H atom for an O descriptor (acsf.py:70-80)


@pytest.mark.parametrize("cutoff", ["tanh", "cos"])
def test_acsf_fp64_fp32_vs_oracle(dev, cutoff):
    from oracle.bridge import sf_rows
    from oracle import pyoracle
    from pantea_amd.bench_models import teacher_acsf

    s = _water("cfg0")
    pos = s.positions.cpu().numpy()
    emap = s.element_map.element_to_atom_type
    for el in ("O", "H"):
        acsf = teacher_acsf(el, cutoff=cutoff)
        _log(f"acsf {el} {cutoff}")
        idx = s.select_host(el)
        ref = pyoracle.acsf(pos, s.atom_types_host, s.box_lengths, sf_rows(acsf, emap), center_pos=pos[idx],
                           self_idx=idx.astype(np.int32))
        g64 = acsf(s).cpu().numpy()
        assert _rel(g64, ref) < 1e-10
        g32 = acsf(s, precision="float32").cpu().numpy()
        assert _rel(g32, ref) < 2e-5


@pytest.mark.parametrize("cutoff", ["tanh", "cos"])
def test_acsf_four_eta_groups_vs_oracle(dev, cutoff):
    """Four eta per angular group spanning 0.002 .. 0.45 (ratios 3, 1.25, 60), both
    cutoff classes, FP64 and FP32 against the oracle."""
    from oracle.bridge import sf_rows
    from oracle import pyoracle
    from pantea_amd.descriptors import ACSF, G2, G3, CutoffFunction, NeighborElements

    s = _water("cfg0")
    pos = s.positions.cpu().numpy()
    emap = s.element_map.element_to_atom_type
    cfn = CutoffFunction.from_type(cutoff, 10.0)
    acsf = ACSF("O", ((G2(cfn, 0.0, 0.02), NeighborElements("H")),),
                tuple((G3(cfn, eta=eta, zeta=z, lambda0=lam), NeighborElements(j, k))
                      for (j, k) in (("H", "H"), ("H", "O")) for eta in (0.002, 0.006, 0.0075, 0.45)
                      for z in (1.0, 4.0) for lam in (1.0, -1.0)))
    idx = s.select_host("O")
    ref = pyoracle.acsf(pos, s.atom_types_host, s.box_lengths, sf_rows(acsf, emap), center_pos=pos[idx],
                       self_idx=idx.astype(np.int32))
    assert np.max(np.abs(ref[:, -8:])) > 0   # the eta = 0.45 terms are not all underflowed
    assert _rel(acsf(s).cpu().numpy(), ref) < 1e-10
    assert _rel(acsf(s, precision="float32").cpu().numpy(), ref) < 2e-5


def test_acsf_grad_vs_oracle(dev):
    from oracle.bridge import sf_rows
    from oracle import pyoracle
    from pantea_amd.bench_models import teacher_acsf

    s = _water("cfg0")
    pos = s.positions.cpu().numpy()
    acsf = teacher_acsf("O")
    rows = sf_rows(acsf, s.element_map.element_to_atom_type)
    idx = np.arange(0, s.natoms, 7)  # atoms of both elements (acsf.py:103-104 semantics)
    J = acsf.grad(s, atom_index=idx).cpu().numpy()
    ref = pyoracle.acsf_selfjac(pos, s.atom_types_host, s.box_lengths, rows, center_pos=pos[idx],
                                self_idx=idx.astype(np.int32))
    assert J.shape == ref.shape
    assert _rel(J, ref) < 1e-10
    J32 = acsf.grad(s, atom_index=idx, precision="float32").cpu().numpy()
    assert _rel(J32, ref) < 1e-4


# ---------------------------------------------------------------------------
# (c) MLP on matrix cores
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("mode,tol", [("float64", 1e-12), ("bf16x3", 2e-4), ("bf16", 3e-2)])
@pytest.mark.parametrize("shape", [(52, (25, 25)), (52, (64, 64)), (34, (16, 16)), (3, (5, 5)), (70, (100, 40, 8))])
def test_mlp_vs_oracle(dev, mode, tol, shape):
    n_in, hidden = shape
    _check_mlp(dev, mode, tol, n_in, hidden, ["tanh", "softplus", "logistic"][: len(hidden)])


@pytest.mark.parametrize("mode,tol", [("float64", 1e-12), ("bf16x3", 2e-4)])
@pytest.mark.parametrize("acts", [("relu", "exp", "identity"), ("gaussian", "cos", "harmonic"), ("tanh", "gaussian")])
def test_mlp_activation_paths_vs_oracle(dev, mode, tol, acts):
    """Activations whose derivative follows from the output (the kernel turns the stored
    output into the derivative in place) and ones whose derivative does not (it keeps a
    derivative buffer per layer), alone and mixed."""
    hidden = (32, 24, 16)[: len(acts)]
    _check_mlp(dev, mode, tol, 40, hidden, list(acts), weight_scale=0.4)


def _check_mlp(dev, mode, tol, n_in, hidden, acts, weight_scale=1.0):
    from oracle import pyoracle
    from pantea_amd import _lib
    from pantea_amd.models import MLP_MODES, DeviceMLP, NeuralNetworkModel, NormalInitializer

    model = NeuralNetworkModel(tuple(zip(hidden, acts)), (1, "identity"), kernel_initializer=NormalInitializer())
    params = model.init(11, n_in)
    for k in params:
        params[k]["kernel"] = params[k]["kernel"] * weight_scale
        params[k]["bias"] = np.random.default_rng(3).normal(size=params[k]["bias"].shape) * 0.1
    n = 203
    rng = np.random.default_rng(5)
    G = rng.random((n, n_in)) * 2.0
    sub, mul = rng.random(n_in), 0.5 + rng.random(n_in)
    add = 0.25
    gdt = torch.float64 if mode == "float64" else torch.float32
    Gd = torch.as_tensor(G, dtype=gdt, device=dev)
    mlp = DeviceMLP(model, params, n_in, np.stack([sub, mul]), add, dev)
    e = torch.zeros(n, dtype=torch.float64, device=dev)
    dE = torch.zeros((n, n_in), dtype=gdt, device=dev)
    mlp.energy_grad(MLP_MODES[mode], n, Gd, n_in, _lib.PT_F64 if mode == "float64" else _lib.PT_F32, e, dE)
    torch.cuda.synchronize()
    layers = [(params[f"layers_{2*i}"]["kernel"], params[f"layers_{2*i}"]["bias"], a)
              for i, a in enumerate(list(acts) + ["identity"])]
    x = (G - sub) * mul + add
    y, g = pyoracle.mlp_forward_backward(layers, x)
    g = g * mul
    assert _rel(e.cpu().numpy(), y) < tol
    assert _rel(dE.cpu().numpy(), g) < tol


# ---------------------------------------------------------------------------
# NNP energy and forces
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def cfg0_teacher(dev):
    from oracle.bridge import oracle_nnp
    from pantea_amd import bench_models

    s = _water("cfg0")
    pot = bench_models.water_nnp(s, hidden=(25, 25), seed=2023)
    onn = oracle_nnp(pot, s)
    ref = onn.energy_forces(s.positions.cpu().numpy(), s.atom_types_host, s.box_lengths)
    return s, pot, onn, ref


def test_nnp_fp64_vs_oracle(cfg0_teacher):
    s, pot, onn, (E, e_atom, f_self, f_full) = cfg0_teacher
    e_gpu, ff = pot.energy_and_forces(s, mode="full")
    _, fs = pot.energy_and_forces(s, mode="reference")
    ea = pot.atomic_energies(s).cpu().numpy()
    _log(f"E gpu={float(e_gpu):.15f} oracle={E:.15f}")
    assert abs(float(e_gpu) - E) <= 1e-10 * np.sum(np.abs(e_atom))
    assert _rel(ea, e_atom) < 1e-10
    assert _rel(fs.cpu().numpy(), f_self) < 1e-10
    assert _rel(ff.cpu().numpy(), f_full) < 1e-10
    assert np.abs(ff.cpu().numpy().sum(0)).max() < 1e-9 * np.abs(f_full).max()
    # compute_forces keeps the reference semantics by default
    assert _rel(pot.compute_forces(s).cpu().numpy(), f_self) < 1e-10


def test_nnp_mixed_precision_error(cfg0_teacher):
    from pantea_amd.units import units

    s, pot, onn, (E, e_atom, f_self, f_full) = cfg0_teacher
    f_rms = np.sqrt(np.mean(f_full ** 2)) * units.TO_MILLI_ELECTRON_VOLT
    # split-BF16 (about 16 mantissa bits per product): the 5 meV/Bohr budget; plain BF16 (8 bits,
    # relative rounding 2^-9 per operand): 1 % of |F| rms and 50 meV/atom
    for prec, f_budget, e_budget in (("mixed", 5.0, 1.0), ("mixed_bf16", 0.01 * f_rms, 50.0)):
        e_m, f_m = pot.energy_and_forces(s, mode="full", precision=prec)
        de = abs(float(e_m) - E) / s.natoms * units.TO_MILLI_ELECTRON_VOLT
        rmse = np.sqrt(np.mean((f_m.cpu().numpy() - f_full) ** 2)) * units.TO_MILLI_ELECTRON_VOLT
        _log(f"{prec}: energy error {de:.4f} meV/atom, force RMSE {rmse:.4f} meV/Bohr "
             f"(|F| rms {f_rms:.1f} meV/Bohr)")
        assert rmse < f_budget and de < e_budget


def test_h2o_fixture_golden_through_api(dev, goldens):
    """test_nnp.py end to end through NeuralNetworkPotential.from_runner + load()."""
    from pantea_amd.io import Dataset
    from pantea_amd.potential import NeuralNetworkPotential

    g = goldens["nnp_outputs"]
    pot = NeuralNetworkPotential.from_runner(GOLDEN / "h2o.json")
    assert pot.num_elements == 2 and pot.elements == ("H", "O")
    pot.load_scaler()
    pot.load_model()
    s = Dataset.from_runner(GOLDEN / "h2o.data")[0]
    F = pot.compute_forces(s).cpu().numpy()
    assert np.allclose(F, np.array(g["forces"]), rtol=1e-5, atol=1e-8)
    assert abs(float(pot(s)) - g["energy"]) < 2.5e-7
    pos, types, L, onn = h2o_fixture_inputs(np.float64)
    E, _, fs, _ = onn.energy_forces(s.positions.cpu().numpy(), s.atom_types_host, s.box_lengths)
    assert _rel(F, fs) < 1e-10


def test_cfg1_subset_and_fd(dev):
    """12,288 atoms: forces on 64 atoms vs oracle (halo-restricted) and vs central FD (h = 1e-5)."""
    from oracle.bridge import oracle_nnp
    from pantea_amd import bench_models
    from pantea_amd.potential import energy

    s = _water("cfg1")
    pot = bench_models.water_nnp(s, hidden=(25, 25), seed=2023)
    rng = np.random.default_rng(64)
    atoms = rng.choice(s.natoms, 64, replace=False)
    x = s.positions.clone().requires_grad_(True)
    e = energy(pot, s, x)
    (grad,) = torch.autograd.grad(e, x)
    f = -grad.cpu().numpy()
    # oracle: only centres within rc of the checked atoms contribute to their forces
    pos = s.positions.cpu().numpy()
    from oracle import pyoracle

    want = np.zeros(s.natoms, np.uint8)
    counts, idx = pyoracle.neighbors(pos, s.box_lengths, 12.0)
    rows = np.split(idx, np.cumsum(counts)[:-1])
    for a in atoms:
        want[a] = 1
        want[rows[a]] = 1
    onn = oracle_nnp(pot, s)
    e_atom, w, _ = onn.atomic(pos, s.atom_types_host, s.box_lengths)
    _, f_full = pyoracle.force_vjp(pos, s.atom_types_host, s.box_lengths, onn._tables_by_type(), w, want)
    assert _rel(f[atoms], f_full[atoms]) < 1e-10
    h = 1e-5
    worst = 0.0
    for a in atoms[:16]:
        for d in range(3):
            p = s.positions.clone(); p[a, d] += h
            m = s.positions.clone(); m[a, d] -= h
            sp, sm = s.copy(), s.copy()
            sp.positions, sm.positions = p, m
            fd = -(float(pot(sp)) - float(pot(sm))) / (2 * h)
            worst = max(worst, abs(fd - f[a, d]) / max(1e-3, abs(f[a, d])))
    _log(f"cfg1 FD worst relative deviation {worst:.2e}")
    assert worst < 1e-5


@pytest.mark.parametrize("cutoff", ["tanh", "tanhu", "cos", "exp", "poly1", "poly2", "hard"])
@pytest.mark.parametrize("zetas,kind", [((1.0, 2.0), "G3"), ((2.0, 3.0), "G3"), ((2.5, 1.0), "G3"), ((1.0, 4.0), "G9"),
                                        ((4.0, 1.0), "G3")])
def test_nnp_cutoff_and_zeta_paths_vs_oracle(dev, cutoff, zetas, kind):
    """Every cutoff type and angular-term pattern the descriptor kernels specialise on
    (zeta all 1 / in {1, 2, 4} / exactly {1, 4} x lambda +-1 / other integers / non-integer
    via pow; G3 and G9), FP64 and mixed, through energy and -grad E against the oracle."""
    from oracle.bridge import oracle_nnp
    from pantea_amd.bench_models import make_nnp
    from pantea_amd.descriptors import ACSF, G2, G3, G9, CutoffFunction, NeighborElements

    cfn = CutoffFunction.from_type(cutoff, 9.0)
    Gk = G3 if kind == "G3" else G9

    def acsf_of(el):
        radial = ((G2(cfn, 0.0, 0.02), NeighborElements("H")), (G2(cfn, 0.5, 0.1), NeighborElements("O")))
        angular = tuple((Gk(cfn, eta=eta, zeta=z, lambda0=lam, r_shift=0.0), NeighborElements(j, k))
                        for (j, k) in (("H", "H"), ("H", "O"), ("O", "O"))
                        for eta in (0.005, 0.03) for z in zetas for lam in (1.0, -1.0))
        return ACSF(el, radial, angular)

    s = _water("cfg0")
    pot = make_nnp(acsf_of, (8, 8), seed=3, precision="float64")
    pot.fit_scaler([s])
    E, e_atom, f_self, f_full = oracle_nnp(pot, s).energy_forces(s.positions.cpu().numpy(), s.atom_types_host,
                                                                 s.box_lengths)
    e64, f64 = pot.energy_and_forces(s, mode="full", precision="float64")
    assert abs(float(e64) - E) <= 1e-10 * np.sum(np.abs(e_atom))
    assert _rel(f64.cpu().numpy(), f_full) < 1e-10
    em, fm = pot.energy_and_forces(s, mode="full", precision="mixed")
    assert abs(float(em) - E) <= 1e-4 * np.sum(np.abs(e_atom))
    assert _rel(fm.cpu().numpy(), f_full) < 3e-3


def test_wrong_kernel_kind_fails_loudly(dev):
    """A table launched with a kernel class that does not cover it yields NaN, not a wrong number."""
    from pantea_amd import _lib
    from pantea_amd.bench_models import teacher_acsf
    from pantea_amd.engine import NeighborList, acsf_forward

    s = _water("cfg0")
    acsf = teacher_acsf("O")
    tab = acsf.device_table(s.element_map, dev)
    assert tab.kind == _lib.PT_ACSF_TANH_Z14
    nl = NeighborList(acsf.r_cutoff, s.element_map.n_types).build(s.positions, s.atom_types, s.box_lengths,
                                                                   s.type_counts())
    t_o = s.element_map.element_to_atom_type["O"]
    b, cnt = nl.type_block(t_o)
    good = acsf_forward(nl, tab, _lib.PT_F64, cnt, center_begin=b).cpu().numpy()
    assert np.isfinite(good).all()
    keep = tab.kind
    try:
        # the wider classes cover this table and agree with the specialised kernel
        for other in (_lib.PT_ACSF_TANH_P4, _lib.PT_ACSF_GENERIC):
            tab.kind = other
            alt = acsf_forward(nl, tab, _lib.PT_F64, cnt, center_begin=b).cpu().numpy()
            assert _rel(alt, good) < 1e-13
        tab.kind = _lib.PT_ACSF_TANH_Z1N2   # this table has 4 terms with zeta = 4: not covered
        bad = acsf_forward(nl, tab, _lib.PT_F64, cnt, center_begin=b).cpu().numpy()
    finally:
        tab.kind = keep
    assert np.isnan(bad).all()
    # and a zeta {1, 2} table is outside the Z14 class
    from pantea_amd.descriptors import ACSF, G2, G3, CutoffFunction, NeighborElements

    cfn = CutoffFunction.from_type("tanh", 9.0)
    p4 = ACSF("O", ((G2(cfn, 0.0, 0.02), NeighborElements("H")),),
              tuple((G3(cfn, eta=0.01, zeta=z, lambda0=lam), NeighborElements("H", "H"))
                    for z in (1.0, 2.0) for lam in (1.0, -1.0)))
    tab2 = p4.device_table(s.element_map, dev)
    assert tab2.kind == _lib.PT_ACSF_TANH_P4
    tab2.kind = _lib.PT_ACSF_TANH_Z14
    nl2 = NeighborList(p4.r_cutoff, s.element_map.n_types).build(s.positions, s.atom_types, s.box_lengths,
                                                                  s.type_counts())
    b2, cnt2 = nl2.type_block(t_o)
    bad2 = acsf_forward(nl2, tab2, _lib.PT_F64, cnt2, center_begin=b2).cpu().numpy()
    assert np.isnan(bad2).all()
