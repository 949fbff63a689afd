"""The MD step without host reads: pt_skin_check, pt_nbr_build_async (gated),
pt_permute_rows_gated / pt_copy_gated, and the gated engine against the
synchronous one.

* the gated build with its gate open produces the synchronous build's list
  bit for bit (including rows longer than the 256-entry stage), and with the
  gate closed leaves every output untouched;
* totals: pair count, longest row, overflow (nothing written past the
  capacity), build counter;
* the skin flag against the host's displacement rule;
* MD trajectories of the gated engine equal those of the synchronous engine
  (FP64 and mixed precision);
* a row bound that a build exceeds is found at the next check and the steps
  since the checkpoint are redone: the trajectory still equals the
  synchronous one.
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _water(n_mol, L, seed=2023):
    from pantea_amd import water
    from pantea_amd.atoms import Structure

    return Structure.from_dict(water.water_box(n_mol, L, seed=seed))


def _sync_and_async(s, rc, gate_value, cap=None, row_limit=1 << 30):
    from pantea_amd import _lib
    from pantea_amd.atoms import host_box
    from pantea_amd.engine import NeighborList

    lib = _lib.load()
    nl = NeighborList(rc, s.element_map.n_types).build(s.positions, s.atom_types, s.box_lengths, s.type_counts())
    n = s.natoms
    dev = s.positions.device
    cap = nl.cap if cap is None else cap
    out = {k: torch.full_like(getattr(nl, k), -7) for k in ("perm", "row_off")}
    out["pos4"] = torch.full_like(nl.pos4, -7.0)
    out["nbr"] = torch.full((cap,), -7, dtype=torch.int32, device=dev)
    out["rev"] = torch.full((cap,), -7, dtype=torch.int32, device=dev)
    gate = torch.tensor([gate_value], dtype=torch.int32, device=dev)
    totals = torch.zeros(4, dtype=torch.int64, device=dev)
    ws = torch.empty(lib.pt_nbr_workspace_bytes(n, s.element_map.n_types), dtype=torch.uint8, device=dev)
    pos = s.positions.to(torch.float64).contiguous()
    box = host_box(s.box_lengths)
    _lib.check(lib.pt_nbr_build_async(n, pos.data_ptr(), 3, s.atom_types.to(torch.int32).contiguous().data_ptr(),
                                      s.element_map.n_types, ctypes.cast(box, ctypes.c_void_p), rc,
                                      out["perm"].data_ptr(), out["pos4"].data_ptr(), out["row_off"].data_ptr(),
                                      out["nbr"].data_ptr(), out["rev"].data_ptr(), cap, gate.data_ptr(),
                                      totals.data_ptr(), row_limit, ws.data_ptr(), ws.numel(), _lib.stream_ptr()),
               "pt_nbr_build_async")
    torch.cuda.synchronize()
    return nl, out, totals.cpu().numpy()


@pytest.mark.parametrize("case", ["cfg0", "long_rows"])
def test_gated_build_equals_sync_build(case):
    if case == "cfg0":
        s, rc = _water(128, 29.56), 12.0
    else:
        # dense clusters: rows longer than the 256-entry stage take the second pass
        from pantea_amd.atoms import Structure

        rng = np.random.default_rng(3)
        pos = np.concatenate([rng.normal(size=(400, 3)) * 1.2 + c for c in (5.0, 20.0)])
        s = Structure.from_dict({"positions": np.mod(pos, 30.0), "elements": ["H"] * 800,
                                 "lattice": np.diag([30.0] * 3)})
        rc = 4.0
    nl, out, tot = _sync_and_async(s, rc, 1)
    n, m = s.natoms, nl.n_pairs
    if case == "long_rows":
        assert nl.max_row > 256
    assert torch.equal(out["perm"], nl.perm) and torch.equal(out["pos4"], nl.pos4)
    assert torch.equal(out["row_off"], nl.row_off)
    assert torch.equal(out["nbr"][:m], nl.nbr[:m]) and torch.equal(out["rev"][:m], nl.rev[:m])
    assert (out["nbr"][m:] == -7).all()
    assert tot.tolist() == [m, nl.max_row, m, 1]


def test_rows_over_the_row_limit_are_reported_without_the_second_pass():
    # the caller reads rows up to 200 entries: the long-row pass is not enqueued, the
    # list stays unwritten and totals[1] tells the caller that it is unusable
    from pantea_amd.atoms import Structure

    rng = np.random.default_rng(3)
    pos = np.concatenate([rng.normal(size=(400, 3)) * 1.2 + c for c in (5.0, 20.0)])
    s = Structure.from_dict({"positions": np.mod(pos, 30.0), "elements": ["H"] * 800,
                             "lattice": np.diag([30.0] * 3)})
    nl, out, tot = _sync_and_async(s, 4.0, 1, row_limit=200)
    assert nl.max_row > 256
    assert tot.tolist() == [nl.n_pairs, nl.max_row, nl.n_pairs, 1] and tot[1] > 200
    assert (out["nbr"] == -7).all() and (out["rev"] == -7).all()
    assert torch.equal(out["row_off"], nl.row_off)


def test_closed_gate_leaves_outputs_untouched():
    s = _water(128, 29.56)
    nl, out, tot = _sync_and_async(s, 12.0, 0)
    for k in ("perm", "row_off", "nbr", "rev"):
        assert (out[k] == -7).all(), k
    assert (out["pos4"] == -7.0).all()
    assert tot.tolist() == [0, 0, 0, 0]


def test_overflow_is_reported_and_nothing_is_written_past_the_capacity():
    s = _water(128, 29.56)
    nl, out, tot = _sync_and_async(s, 12.0, 1, cap=1000)
    assert tot[2] == nl.n_pairs > 1000 and tot[3] == 1
    assert (out["nbr"] == -7).all() and (out["rev"] == -7).all()


def test_skin_flag_matches_the_displacement_rule():
    from pantea_amd import _lib
    from pantea_amd.atoms import host_box

    lib = _lib.load()
    rng = np.random.default_rng(1)
    n, L = 5000, 40.0
    ref = np.zeros((n, 4))
    ref[:, :3] = rng.random((n, 3)) * L
    for d1, d2, skin in ((0.3, 0.2, 1.0), (0.6, 0.5, 1.0), (0.5, 0.5, 1.0), (0.1, 0.0, 0.0)):
        pos = ref.copy()
        pos[17, 0] += d1
        pos[4242, 1] -= d2
        pos[:, :3] = np.mod(pos[:, :3], L)
        a = torch.as_tensor(pos, device="cuda")
        b = torch.as_tensor(ref, device="cuda")
        scratch = torch.zeros(514, dtype=torch.float64, device="cuda")
        flag = torch.full((1,), -1, dtype=torch.int32, device="cuda")
        box = host_box((L, L, L))
        _lib.check(lib.pt_skin_check(n, a.data_ptr(), b.data_ptr(), ctypes.cast(box, ctypes.c_void_p), skin,
                                     scratch.data_ptr(), flag.data_ptr(), _lib.stream_ptr()), "pt_skin_check")
        want = 1 if (skin <= 0.0 or d1 + d2 > skin) else 0
        assert int(flag.item()) == want, (d1, d2, skin)


@pytest.mark.parametrize("gate_value", [0, 1])
def test_gated_permutation(gate_value):
    from pantea_amd import _lib

    lib = _lib.load()
    n = 1000
    perm = torch.as_tensor(np.random.default_rng(2).permutation(n).astype(np.int32), device="cuda")
    gate = torch.tensor([gate_value], dtype=torch.int32, device="cuda")
    for width, dt in ((3, torch.float64), (1, torch.float64), (1, torch.int32)):
        buf = (torch.arange(n * width, device="cuda").reshape(n, width)).to(dt)
        before = buf.clone()
        tmp = torch.empty_like(buf)
        _lib.check(lib.pt_permute_rows_gated(n, width, buf.element_size(), perm.data_ptr(), buf.data_ptr(),
                                             tmp.data_ptr(), gate.data_ptr(), _lib.stream_ptr()), "permute")
        torch.cuda.synchronize()
        want = before[perm.long()] if gate_value else before
        assert torch.equal(buf, want)


@pytest.mark.parametrize("gate_value", [0, 1])
def test_fused_state_permutation(gate_value):
    """pt_permute_state_gated moves several arrays (FP64 and int32, widths 1 and 3) by one
    permutation in two launches; closed gate: nothing moves."""
    from pantea_amd import _lib

    lib = _lib.load()
    n = 1237
    perm = torch.as_tensor(np.random.default_rng(5).permutation(n).astype(np.int32), device="cuda")
    gate = torch.tensor([gate_value], dtype=torch.int32, device="cuda")
    bufs = [torch.randn((n, 3), dtype=torch.float64, device="cuda"), torch.randn(n, dtype=torch.float64, device="cuda"),
            torch.arange(n, dtype=torch.int32, device="cuda") * 7, torch.randint(0, 9, (n, 3), dtype=torch.int32,
                                                                                 device="cuda")]
    before = [b.clone() for b in bufs]
    arrays = (_lib.PermuteArray * 4)(*[_lib.PermuteArray(b.data_ptr(), 1 if b.dim() == 1 else b.shape[1],
                                                         b.element_size()) for b in bufs])
    nbytes = lib.pt_permute_state_tmp_bytes(n, 4, arrays)
    assert nbytes == n * (24 + 8 + 4 + 12)
    tmp = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    _lib.check(lib.pt_permute_state_gated(n, perm.data_ptr(), 4, arrays, tmp.data_ptr(), tmp.numel(), gate.data_ptr(),
                                          _lib.stream_ptr()), "pt_permute_state_gated")
    torch.cuda.synchronize()
    for b, b0 in zip(bufs, before):
        assert torch.equal(b, b0[perm.long()] if gate_value else b0)
    # an int32 array ahead of an FP64 one with an odd count would misalign the FP64 rows: refused
    odd = (_lib.PermuteArray * 2)(_lib.PermuteArray(bufs[2].data_ptr(), 1, 4), _lib.PermuteArray(bufs[1].data_ptr(), 1, 8))
    assert lib.pt_permute_state_gated(n, perm.data_ptr(), 2, odd, tmp.data_ptr(), tmp.numel(), gate.data_ptr(),
                                      _lib.stream_ptr()) != 0


def _trajectory(s, pot, precision, sync_free, steps=60, row_bound=None, thermostat=True, skin=1.0, graph=False):
    from pantea_amd.simulation import BrendsenThermostat, MDSimulator, System
    from pantea_amd.units import units

    sysm = System.from_structure(s, pot, temperature=300.0, seed=2023, force_mode="full", precision=precision,
                                 skin=skin, sync_free=sync_free, graph=graph)
    md = sysm._engine()
    if row_bound is not None:
        md.row_bound = md.nl.max_row + row_bound
    dt = 0.5 * units.FROM_FEMTO_SECOND
    sim = MDSimulator(dt, thermostat=BrendsenThermostat(300.0, dt) if thermostat else None, integrator="physical")
    sim.run(sysm, steps)
    x, v = sysm.positions.cpu().numpy(), sysm.velocities.cpu().numpy()
    return x, v, md.n_builds, md.redos


@pytest.mark.parametrize("skin", [1.0, 0.0])
@pytest.mark.parametrize("precision", ["float64", "mixed"])
def test_gated_md_equals_synchronous_md(precision, skin):
    # skin 0: no displacement check, a rebuild every step
    from pantea_amd.bench_models import water_nnp

    s = _water(128, 29.56)
    pot = water_nnp(s, hidden=(25, 25), seed=2023)
    x0, v0, b0, _ = _trajectory(s, pot, precision, False, skin=skin)
    x1, v1, b1, _ = _trajectory(s, pot, precision, True, skin=skin)
    print(f"{precision}, skin {skin}: {b0} synchronous builds, {b1} gated builds", flush=True)
    assert b0 == b1 and b0 > 2
    if skin == 0.0:
        assert b0 == 61
    np.testing.assert_array_equal(x1, x0)
    np.testing.assert_array_equal(v1, v0)


def test_row_bound_overflow_is_redone():
    """A row bound below the real longest row makes the descriptor kernels write NaN for the
    long rows; the check of the next read finds the rebuilds that exceeded it and the steps
    since the checkpoint are redone with a larger bound."""
    from pantea_amd.bench_models import water_nnp

    s = _water(128, 29.56)
    pot = water_nnp(s, hidden=(25, 25), seed=2023)
    x0, v0, _, _ = _trajectory(s, pot, "float64", False, steps=30, skin=0.0)
    # without a skin every step rebuilds; a bound a few entries below the longest row of
    # the start is exceeded by the first rebuild
    x1, v1, _, redos = _trajectory(s, pot, "float64", True, steps=30, row_bound=-8, skin=0.0)
    assert redos >= 1
    assert np.isfinite(x1).all()
    np.testing.assert_array_equal(x1, x0)
    np.testing.assert_array_equal(v1, v0)


@pytest.mark.parametrize("precision", ["float64", "mixed"])
def test_graph_replayed_md_equals_synchronous_md(precision):
    """The gated step captured once into a HIP graph and replayed (recaptured after each
    checkpoint restore and buffer change) gives the synchronous trajectory bit for bit."""
    from pantea_amd.bench_models import water_nnp

    s = _water(128, 29.56)
    pot = water_nnp(s, hidden=(25, 25), seed=2023)
    x0, v0, b0, _ = _trajectory(s, pot, precision, False)
    x1, v1, b1, _ = _trajectory(s, pot, precision, True, graph=True)
    assert b0 == b1
    np.testing.assert_array_equal(x1, x0)
    np.testing.assert_array_equal(v1, v0)
