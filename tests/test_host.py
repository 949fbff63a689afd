"""Host-side logic and the C ABI library, on the CPU (no kernel launches)."""

import ctypes
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import GOLDEN, ROOT


def test_library_exports_every_declared_symbol():
    from pantea_amd import _lib

    lib = _lib.load()
    header = (ROOT / "include" / "pantea_hip.h").read_text()
    declared = set()
    import re

    for m in re.finditer(r"^\s*(?:const\s+)?[a-z_]+\s*\*?\s*(pt_[a-z0-9_]+)\s*\(", header, re.M):
        declared.add(m.group(1))
    assert declared, "no declarations parsed"
    assert declared == set(_lib.EXPORTED_SYMBOLS), declared ^ set(_lib.EXPORTED_SYMBOLS)
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.pt_version() >= 1
    assert lib.pt_nbr_workspace_bytes(1000, 2) > 0


def test_struct_layouts_match_header():
    from pantea_amd import _lib

    # sizes as laid out by the C compiler
    src = r'''
    #include <stdio.h>
    #include "pantea_hip.h"
    int main(void){printf("%zu %zu %zu %zu %zu\n", sizeof(pt_radial_group), sizeof(pt_angular_group),
    sizeof(pt_sf_table), sizeof(pt_mlp_desc), sizeof(pt_permute_array)); return 0;}
    '''
    tmp = Path("/tmp/pt_sizes.c")
    tmp.write_text(src)
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(tmp), "-o", "/tmp/pt_sizes"], check=True)
    out = subprocess.run(["/tmp/pt_sizes"], capture_output=True, text=True, check=True).stdout.split()
    assert [int(x) for x in out] == [ctypes.sizeof(_lib.RadialGroup), ctypes.sizeof(_lib.AngularGroup),
                                     ctypes.sizeof(_lib.SFTable), ctypes.sizeof(_lib.MlpDesc),
                                     ctypes.sizeof(_lib.PermuteArray)]


def test_product_fails_loudly_without_gpu():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from pantea_amd import _lib
    from pantea_amd.atoms import Structure

    with pytest.raises(_lib.HipLibraryError):
        Structure.from_dict({"positions": [[0, 0, 0]], "elements": ["H"]})


def test_table_compiler_groups_and_columns():
    from pantea_amd.bench_models import teacher_acsf
    from pantea_amd.engine import compile_table

    acsf = teacher_acsf("O")
    rows = acsf.rows({"H": 1, "O": 2})
    assert len(rows) == 52 and [r.col for r in rows] == list(range(52))
    t = compile_table(rows)
    assert t.n_sf == 52 and t.n_rg == 2 and t.n_ag == 3
    cols = set()
    for g in t.rg[: t.n_rg]:
        assert g.n == 8
        cols |= set(g.col[: g.n])
    for g in t.ag[: t.n_ag]:
        assert (g.n_eta, g.n_ang) == (3, 4)
        assert list(g.pre[:4]) == [1.0, 1.0, 0.125, 0.125]
        assert list(g.izeta[:4]) == [1, 1, 4, 4]
        cols |= {c for c in g.col if c >= 0}
    assert cols == set(range(52))


def test_table_compiler_splits_oversized_groups():
    from pantea_amd.descriptors import ACSF, G3, CutoffFunction, NeighborElements
    from pantea_amd.engine import compile_table

    cfn = CutoffFunction.from_type("cos", 6.0)
    ang = tuple((G3(cfn, eta=eta, zeta=z, lambda0=1.0), NeighborElements("H", "H"))
                for eta in (0.1, 0.2, 0.3, 0.4, 0.5, 0.6) for z in (1.0, 2.0, 3.0, 4.5, 5.0))
    t = compile_table(ACSF("H", (), ang).rows({"H": 1}))
    seen = []
    for g in t.ag[: t.n_ag]:
        seen += [c for c in g.col if c >= 0]
        for a in range(g.n_ang):
            if g.zeta[a] == 4.5:
                assert g.izeta[a] == -1
    assert sorted(seen) == list(range(30))


def test_table_kind_classification():
    """pt_sf_table_kind (host code) picks the specialised descriptor kernels only for tables they cover."""
    import ctypes

    from pantea_amd import _lib
    from pantea_amd.bench_models import student_acsf, teacher_acsf
    from pantea_amd.descriptors import ACSF, G2, G3, CutoffFunction, NeighborElements
    from pantea_amd.engine import compile_table

    lib = _lib.load()
    emap = {"H": 1, "O": 2}

    def kind(acsf):
        return lib.pt_sf_table_kind(ctypes.byref(compile_table(acsf.rows(emap))))

    assert kind(teacher_acsf("O")) == _lib.PT_ACSF_TANH_Z14         # zeta {1, 4} x lambda {+1, -1}, tanh
    assert kind(student_acsf("H")) == _lib.PT_ACSF_TANH_Z1N2        # zeta = 1, 2 terms, tanh
    assert kind(teacher_acsf("O", cutoff="tanhu")) == _lib.PT_ACSF_TANH_Z14
    assert kind(teacher_acsf("O", cutoff="cos")) == _lib.PT_ACSF_GENERIC
    cfn = CutoffFunction.from_type("tanh", 6.0)

    def angular(terms):
        return ACSF("H", ((G2(cfn, 0.0, 0.1), NeighborElements("H")),),
                    tuple((G3(cfn, eta=eta, zeta=z, lambda0=lam), NeighborElements("H", "O"))
                          for eta in (0.01, 0.1) for z, lam in terms))

    z14 = ((1.0, 1.0), (1.0, -1.0), (4.0, 1.0), (4.0, -1.0))
    assert kind(angular(z14)) == _lib.PT_ACSF_TANH_Z14
    assert kind(angular(z14[::-1])) == _lib.PT_ACSF_TANH_Z14           # any order of the four terms
    # three of them: the compiler completes the group with a term that owns no column
    assert kind(angular(z14[:3])) == _lib.PT_ACSF_TANH_Z14
    t = compile_table(angular(z14[:3]).rows(emap))
    g = t.ag[0]
    assert g.n_ang == 4 and t.n_sf == 7
    assert sorted(c for c in g.col if c >= 0) == list(range(1, 7))
    assert [g.col[e * _lib.PT_AG_ANG + 3] for e in range(2)] == [-1, -1]
    assert (g.zeta[3], g.lam[3]) == (4.0, -1.0)
    # not when a group has another term, nor for the cheaper zeta = 1 class
    assert kind(angular(z14[:2])) == _lib.PT_ACSF_TANH_Z1N2
    assert compile_table(angular(z14[:2]).rows(emap)).ag[0].n_ang == 2
    assert kind(angular(((1.0, 1.0), (2.0, -1.0), (4.0, 1.0), (4.0, -1.0)))) == _lib.PT_ACSF_TANH_P4
    z3 = ACSF("H", ((G2(cfn, 0.0, 0.1), NeighborElements("H")),),
              ((G3(cfn, eta=0.1, zeta=3.0, lambda0=1.0), NeighborElements("H", "H")),))
    assert kind(z3) == _lib.PT_ACSF_GENERIC
    radial_only = ACSF("H", ((G2(cfn, 0.0, 0.1), NeighborElements("H")),), ())
    assert kind(radial_only) == _lib.PT_ACSF_GENERIC


def test_settings_from_json_fixture():
    from pantea_amd.potential import NNPSettings

    s = NNPSettings.from_file(GOLDEN / "h2o.json")
    assert s.elements == ["H", "O"] and s.cutoff_type == "tanhu" and s.scale_type == "center"
    assert s.global_nodes_short == [5, 5] and s.global_activation_short == ["tanh", "tanh", "identity"]
    assert len(s.symfunction_short) == 7
    a = s.symfunction_short[-1]
    assert (a.acsf_type, a.neighbor_element_j, a.neighbor_element_k, a.lambda0, a.zeta) == (3, "O", "O", -1.0, 4.0)


def test_settings_from_nn(tmp_path):
    from pantea_amd.potential import NNPSettings

    p = tmp_path / "input.nn"
    p.write_text(
        "# comment\nnumber_of_elements 2\nelements O H\ncutoff_type 1\nscale_symmetry_functions\n"
        "global_nodes_short 10 10\nglobal_activation_short t t l\n"
        "symfunction_short H 2 O 0.01 0.0 12.0\n"
        "symfunction_short O 3 H H 0.02 -1.0 4.0 12.0\n")
    s = NNPSettings.from_file(p)
    assert s.elements == ["H", "O"] and s.cutoff_type == "cos" and s.scale_type == "scale"
    r, a = s.symfunction_short
    assert (r.eta, r.r_shift, r.r_cutoff) == (0.01, 0.0, 12.0)
    assert (a.eta, a.lambda0, a.zeta, a.r_cutoff) == (0.02, -1.0, 4.0, 12.0)
    # trainer keywords keep the reference's defaults (settings.py:106-131)
    assert (s.epochs, s.updater_type, s.kalman_type, s.kalman_epsilon, s.kalman_lambda_short) == \
        (1, "gradient_descent", 0, 0.01, 0.96)


def test_settings_trainer_keywords(tmp_path):
    from pantea_amd.potential import NNPSettings

    p = tmp_path / "input.nn"
    p.write_text(
        "elements H O\nglobal_nodes_short 5 5\nglobal_activation_short t t l\n"
        "epochs 7\nupdater_type 1\nforce_weight 3.5\nshort_force_fraction 0.25\nshort_energy_fraction 0.75\n"
        "kalman_type 1\nkalman_epsilon 0.02\nkalman_q0 0.001\nkalman_qtau 1.5\nkalman_qmin 1e-7\n"
        "kalman_eta 0.05\nkalman_etatau 2.0\nkalman_etamax 0.5\nkalman_lambda_short 0.98\n"
        "kalman_neu_short 0.999\n")
    s = NNPSettings.from_file(p)
    assert (s.epochs, s.updater_type, s.force_weight, s.short_force_fraction, s.short_energy_fraction) == \
        (7, "kalman_filter", 3.5, 0.25, 0.75)
    assert (s.kalman_type, s.kalman_epsilon, s.kalman_q0, s.kalman_qtau, s.kalman_qmin, s.kalman_eta,
            s.kalman_etatau, s.kalman_etamax, s.kalman_lambda_short, s.kalman_neu_short) == \
        (1, 0.02, 0.001, 1.5, 1e-7, 0.05, 2.0, 0.5, 0.98, 0.999)
    j = NNPSettings.from_file(GOLDEN / "h2o.json")
    assert (j.epochs, j.updater_type, j.force_weight) == (20, "gradient_descent", 10.0)


def test_runner_reader():
    from pantea_amd.io import read_runner_blocks

    blocks = list(read_runner_blocks(GOLDEN / "h2o.data"))
    assert len(blocks) == 2
    b = blocks[0]
    assert len(b["positions"]) == 12 and b["elements"][:3] == ["O", "H", "H"]
    assert b["total_energy"] == [-32.1390027258]


def test_water_generator_geometry():
    from pantea_amd import water

    d = water.water_box(64, water.box_for_density(64), seed=1, d_min=4.5)
    pos = d["positions"]
    L = d["lattice"][0, 0]
    assert pos.min() >= 0 and pos.max() < L
    O_, H1, H2 = pos[0::3], pos[1::3], pos[2::3]

    def mic(v):
        return v - L * np.round(v / L)

    r1 = np.linalg.norm(mic(H1 - O_), axis=1)
    r2 = np.linalg.norm(mic(H2 - O_), axis=1)
    assert abs(np.mean(np.r_[r1, r2]) - 1.81) < 0.02
    cos = np.sum(mic(H1 - O_) * mic(H2 - O_), axis=1) / (r1 * r2)
    assert abs(np.degrees(np.mean(np.arccos(cos))) - 104.5) < 2.0
    dmin = min(np.linalg.norm(mic(O_[i] - O_[j])) for i in range(64) for j in range(i))
    assert dmin > 4.5 - 4 * 0.05 * np.sqrt(2) * 3
    # 1 g/cm^3: 64 molecules in (L Bohr)^3
    assert water.box_for_density(128) == pytest.approx(29.56, abs=0.1)


def test_units_and_masses(goldens):
    from pantea_amd.atoms import ElementMap
    from pantea_amd.units import units

    for e in ("H", "O", "Ne"):
        assert ElementMap.get_atomic_mass_from_element(e) == pytest.approx(goldens["masses"][e], rel=1e-15)
    assert units.FROM_FEMTO_SECOND * 0.5 == pytest.approx(20.670687, rel=1e-6)
    em = ElementMap.from_list(["O", "H", "H"])
    assert em.element_to_atom_type == {"H": 1, "O": 2} and em.unique_elements == ("H", "O")


def test_trainer_settings_and_updater_choice():
    """trainer.py:29-44 (TrainingParams) and :150-171 (updater by updater_type)."""
    from pantea_amd.kalman import NeuralNetworkPotentialTrainer, TrainingParams
    from pantea_amd.potential import NNPSettings, NeuralNetworkPotential

    s = NNPSettings.from_file(GOLDEN / "h2o.json")
    tp = TrainingParams.from_runner(s)
    assert (tp.force_weight, tp.energy_fraction, tp.force_fraction, tp.epochs) == (10.0, 1.0, 0.1, 20)
    pot = NeuralNetworkPotential.from_settings(s, directory=GOLDEN)
    with pytest.raises(NotImplementedError, match="gradient_descent"):
        NeuralNetworkPotentialTrainer.from_settings(pot, s)   # h2o.json asks for gradient descent
