"""Scaler statistics and the extrapolation counter on the GPU
(pt_scaler_partial_fit, pt_scaler_extrapolation) against the oracle's
closed forms (oracle/pyoracle.py scaler_fit / scaler_partial_fit, which
restate descriptors/scaler.py:249-283) and numpy counts.
"""

import numpy as np
import pytest

from oracle import pyoracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _rows(rng, n, d, dtype):
    x = rng.normal(size=(n, d)) * rng.uniform(0.1, 10.0, size=d) + rng.normal(size=d) * 5.0
    return x.astype(dtype).astype(np.float64)


@pytest.mark.parametrize("dtype,n", [(np.float64, 98_304), (np.float32, 5_000), (np.float64, 1)])
def test_device_fit_matches_oracle(dtype, n):
    from pantea_amd.descriptors import DescriptorScaler

    rng = np.random.default_rng(n)
    d = 52
    x = _rows(rng, n, d, dtype)
    # rows with a leading dimension wider than the width (as the descriptor buffers have)
    buf = torch.zeros((n, 64), dtype=torch.float64 if dtype == np.float64 else torch.float32, device="cuda")
    buf[:, :d] = torch.as_tensor(x, device="cuda")
    p = DescriptorScaler.fit(buf[:, :d])
    o = pyoracle.scaler_fit(x)
    assert p.nsamples == n and p.dimension == d
    np.testing.assert_allclose(p.mean, o["mean"], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(p.sigma, o["sigma"], rtol=1e-11, atol=1e-13)
    assert np.array_equal(p.minval, o["minval"]) and np.array_equal(p.maxval, o["maxval"])


def test_device_partial_fit_matches_oracle_chunks():
    from pantea_amd.descriptors import DescriptorScaler

    rng = np.random.default_rng(3)
    x = _rows(rng, 20_000, 34, np.float64)
    cuts = (0, 700, 701, 9_000, 20_000)
    dev = torch.as_tensor(x, device="cuda")
    p = DescriptorScaler.fit(dev[cuts[0]:cuts[1]])
    o = pyoracle.scaler_fit(x[cuts[0]:cuts[1]])
    for a, b in zip(cuts[1:-1], cuts[2:]):
        p = DescriptorScaler.partial_fit(p, dev[a:b])
        o = pyoracle.scaler_partial_fit(o, x[a:b])
    assert p.nsamples == o["nsamples"] == 20_000
    np.testing.assert_allclose(p.mean, o["mean"], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(p.sigma, o["sigma"], rtol=1e-12, atol=1e-13)
    assert np.array_equal(p.minval, o["minval"]) and np.array_equal(p.maxval, o["maxval"])
    # and the merged statistics are those of the whole set (closed form)
    np.testing.assert_allclose(p.mean, x.mean(0), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(p.sigma, x.std(0), rtol=1e-10, atol=1e-12)


def test_extrapolation_counter_matches_numpy():
    from pantea_amd.descriptors import DescriptorScaler, extrapolation_counts, params_to_device

    rng = np.random.default_rng(5)
    fit = _rows(rng, 4_000, 52, np.float64)
    p = DescriptorScaler.fit(torch.as_tensor(fit, device="cuda"))
    new = _rows(rng, 3_000, 52, np.float64) * 1.15
    lo, hi = p.minval, p.maxval
    w = hi - lo
    dev_p = params_to_device(p, "cuda")
    # the second launch reads FP32 rows: the expected counts use the same rounded values
    new32 = new.copy()
    new32[1000:] = new[1000:].astype(np.float32).astype(np.float64)
    excess_all = np.maximum(new32 - hi, lo - new32) / w
    for margin in (0.0, 0.1):
        out = ~((new32 >= lo - margin * w) & (new32 <= hi + margin * w))
        c = torch.zeros(3, dtype=torch.int64, device="cuda")
        extrapolation_counts(torch.as_tensor(new[:1000], device="cuda"), dev_p, c, margin)   # launches accumulate
        extrapolation_counts(torch.as_tensor(new[1000:], device="cuda").float(), dev_p, c, margin)
        got = c.cpu().numpy()
        assert got[0] == np.count_nonzero(out.any(1)) > 0
        assert got[1] == np.count_nonzero(out)
        excess = torch.tensor([int(got[2])], dtype=torch.int64).view(torch.float64).item()
        assert float(np.max(excess_all)) > margin
        assert excess == pytest.approx(float(np.max(excess_all)), rel=1e-12)
    # the rows the scaler was fitted on never extrapolate
    c = torch.zeros(3, dtype=torch.int64, device="cuda")
    extrapolation_counts(torch.as_tensor(fit, device="cuda"), dev_p, c, 0.0)
    assert c.cpu().numpy().tolist() == [0, 0, 0]
    # a NaN is out of range, with an infinite excess
    bad = fit[:10].copy()
    bad[4, 7] = np.nan
    c.zero_()
    extrapolation_counts(torch.as_tensor(bad, device="cuda"), dev_p, c, 0.0)
    got = c.cpu().numpy()
    assert got[0] == 1 and got[1] == 1
    assert torch.tensor([int(got[2])], dtype=torch.int64).view(torch.float64).item() == float("inf")


def test_check_warnings_on_device_rows(capsys):
    """check_warnings (scaler.py:112-142): one warning per call with any value out of range."""
    from pantea_amd.descriptors import DescriptorScaler

    rng = np.random.default_rng(9)
    fit = torch.as_tensor(_rows(rng, 500, 8, np.float64), device="cuda")
    p = DescriptorScaler.fit(fit)
    w = DescriptorScaler.initialize_warnings(0, 2)
    w = DescriptorScaler.check_warnings(p, fit, w)
    assert w.number_of_warnings == 0
    w = DescriptorScaler.check_warnings(p, fit * 2.0 + 1.0, w)
    assert w.number_of_warnings == 1
    off = DescriptorScaler.initialize_warnings()   # max < 0: not counted
    assert DescriptorScaler.check_warnings(p, fit * 5.0, off).number_of_warnings == 0


def test_potential_scaler_fit_on_device_matches_oracle():
    """NeuralNetworkPotential.fit_scaler fits on GPU descriptors with the device kernels: the same
    statistics as the oracle's fit of the oracle's FP64 descriptors."""
    from oracle.bridge import sf_rows
    from pantea_amd import water
    from pantea_amd.atoms import Structure
    from pantea_amd.bench_models import make_nnp, teacher_acsf

    s = Structure.from_dict(water.water_box(128, 29.56, seed=2023))
    pot = make_nnp(teacher_acsf, (8, 8), seed=1)
    pot.fit_scaler([s])
    pos = s.positions.cpu().numpy()
    types = s.atom_types_host
    for e in pot.elements:
        t = s.element_map.element_to_atom_type[e]
        idx = np.nonzero(types == t)[0]
        G = pyoracle.acsf(pos, types, s.box_lengths, sf_rows(pot.atomic_potentials[e].descriptor,
                                                             s.element_map.element_to_atom_type),
                          center_pos=pos[idx], self_idx=idx.astype(np.int32))
        o = pyoracle.scaler_fit(G)
        p = pot.scalers_params[e]
        np.testing.assert_allclose(p.mean, o["mean"], rtol=1e-10, atol=1e-14)
        np.testing.assert_allclose(p.sigma, o["sigma"], rtol=1e-8, atol=1e-14)
        np.testing.assert_allclose(p.minval, o["minval"], rtol=1e-10, atol=1e-14)
        np.testing.assert_allclose(p.maxval, o["maxval"], rtol=1e-10, atol=1e-14)
