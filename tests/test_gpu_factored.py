"""GPU parity tests of the band-sum-first FP64 path (precision
"fp64_factored"): P = sum_n1 conj(aqsm) * aqsn on the FP64 matrix cores, then
sum P * delw * I_eps.  Same bar as the per-term FP64 kernel: within 1e-11
relative of the oracle per achtemp/asxtemp entry, on ragged shapes (bands not
a multiple of the 4-band matrix step, G' rows not a multiple of 64, ncouls
not a multiple of the 64 G-vector tile), skipped G' rows, exact resonances,
every semantics, host and device inputs.
"""
import numpy as np
import pytest

from conftest import gppk, gpp_oracle as orc, rel_err
from test_gpu_parity import FP64_TOL, check, say, _sweep_shapes
from test_oracle_cpu import closed_form

pytestmark = pytest.mark.gpu

FACT = "fp64_factored"


@pytest.mark.parametrize("shape", [(64, 2, 2048, 8), (8, 2, 64, 4), (5, 1, 40, 8),
                                   (12, 0, 100, 7), (9, 3, 1000, 13), (7, 0, 129, 1)])
@pytest.mark.parametrize("sem", ["gppker", "strict", "bgw"])
def test_factored_constfill_closed_form(handle, shape, sem):
    nb, nv, nc, npg = shape
    say("factored constfill", shape, sem)
    inp = gppk.fill_inputs(1, nb, nc, npg, nvband=nv)
    got = handle.sigma(inp, precision=FACT, semantics=sem)
    ach, asx = closed_form(nb, nv, nc, npg, "bgw" if sem == "bgw" else "gppker")
    assert rel_err(got["achtemp"], ach) < 1e-12
    if min(nv, nb) > 0:
        assert rel_err(got["asxtemp"], asx) < 1e-12


@pytest.mark.parametrize("shape", [(6, 2, 256, 4), (5, 1, 1000, 10), (4, 0, 777, 7),
                                   (7, 2, 4096, 32), (11, 2, 300, 3), (2, 1, 130, 2),
                                   (35, 2, 2200, 10)])
@pytest.mark.parametrize("sem", ["gppker", "strict", "bgw"])
def test_factored_random_vs_oracle(handle, shape, sem):
    nb, nv, nc, npg = shape
    say("factored random", shape, sem)
    inp = gppk.fill_inputs(2, nb, nc, npg, nvband=nv, seed=nb * 31 + nc)
    got = handle.sigma(inp, precision=FACT, semantics=sem)
    ref = orc.sigma(inp, sem)
    check(got, ref)
    assert abs(got["achstemp"] - ref["achstemp"]) <= 1e-11 * abs(ref["achstemp"])


@pytest.mark.parametrize("shape", _sweep_shapes(16, 2025))
def test_factored_shape_sweep_vs_oracle(handle, shape):
    nb, nv, nc, npg = shape
    say("factored sweep", shape)
    inp = gppk.fill_inputs(2, nb, nc, npg, nvband=nv, seed=nb + 7 * nc)
    for sem in ("gppker", "strict", "bgw"):
        check(handle.sigma(inp, precision=FACT, semantics=sem), orc.sigma(inp, sem))


@pytest.mark.parametrize("sem", ["gppker", "bgw"])
def test_factored_nearly_real_band_products(handle, sem):
    """The band product forms each complex product from three real ones
    ((Ar - Ai)(Br + Bi) among them), whose error is bounded normwise: with
    aqsm and aqsn nearly real (imaginary parts 1e-6 of the real ones) Im P
    is tiny beside Re P, the case where that form loses most against the
    four-product one.  Every achtemp/asxtemp entry still within 1e-11 of the
    oracle and of the per-term FP64 path."""
    say("factored nearly real", sem)
    inp = gppk.fill_inputs(2, 40, 2000, 8, nvband=2, seed=77)
    q = dict(inp)
    for k in ("aqsmtemp", "aqsntemp"):
        a = inp[k]
        q[k] = np.ascontiguousarray(a.real + 1e-6j * a.imag)
    got = handle.sigma(q, precision=FACT, semantics=sem)
    check(got, orc.sigma(q, sem))
    check(got, handle.sigma(q, semantics=sem))


def test_factored_all_bands_unocc(handle):
    """achtemp_kernel semantics: every band non-occupied."""
    say("factored all bands unocc")
    inp = gppk.fill_inputs(2, 10, 900, 6, nvband=3, seed=77)
    got = handle.sigma(inp, precision=FACT, flags=gppk.FLAG_ALL_BANDS_UNOCC)
    check(got, orc.sigma(inp, flags=orc.ALL_BANDS_UNOCC), keys=("achtemp",))


def test_factored_skipped_rows_and_index_edge(handle):
    say("factored index edges")
    inp = gppk.fill_inputs(2, 6, 320, 4, nvband=1, seed=3)
    inv = inp["inv_igp_index"].copy()
    inv[1] = -5
    inv[2] = 320
    inv[3] = 10_000
    q = dict(inp, inv_igp_index=inv)
    check(handle.sigma(q, precision=FACT), orc.sigma(q))


def test_factored_exact_resonance_nan_pattern(handle):
    say("factored resonance")
    inp = gppk.fill_inputs(2, 3, 256, 4, nvband=0, seed=5)
    w = inp["wtilde_array"].copy()
    w[0, 5] = 4.0 + 0.0j
    q = dict(inp, wtilde_array=w)
    got = handle.sigma(q, precision=FACT)
    ref = orc.sigma(q)
    assert np.isnan(got["achtemp"][1]) and np.isnan(ref["achtemp"][1])
    assert rel_err(got["achtemp"][[0, 2]], ref["achtemp"][[0, 2]]) <= FP64_TOL
    check(handle.sigma(q, precision=FACT, semantics="bgw"), orc.sigma(q, "bgw"))


def test_factored_linearity(handle):
    """achtemp is linear in aqsn and conjugate-linear in aqsm: scaling aqsn
    by 2 and aqsm by i scales the non-occupied achtemp by 2 * conj(i)."""
    say("factored linearity")
    inp = gppk.fill_inputs(2, 9, 1500, 6, nvband=0, seed=19)
    a = handle.sigma(inp, precision=FACT, flags=gppk.FLAG_NO_ACHSTEMP)
    q = dict(inp, aqsntemp=2.0 * inp["aqsntemp"], aqsmtemp=1j * inp["aqsmtemp"])
    b = handle.sigma(q, precision=FACT, flags=gppk.FLAG_NO_ACHSTEMP)
    assert rel_err(b["achtemp"], -2j * a["achtemp"]) <= 1e-13


def test_factored_host_pipeline_matches_device(handle):
    """Host inputs: chunked G' rows, one launch per chunk, same records as
    the single launch, so the same bits as device inputs."""
    import torch
    say("factored host pipeline")
    inp = gppk.fill_inputs(2, 9, 8000, 4, nvband=2, seed=91)
    dev = {k: (torch.from_numpy(np.ascontiguousarray(v)).cuda() if isinstance(v, np.ndarray) else v)
           for k, v in inp.items()}
    torch.cuda.synchronize()
    a = handle.sigma(inp, precision=FACT)
    b = handle.sigma(dev, precision=FACT, memory=gppk.MEM_DEVICE)
    for k in ("achtemp", "asxtemp"):
        assert np.array_equal(a[k], b[k]), (k, a[k], b[k])
    check(a, orc.sigma(inp))


def test_factored_canonical_shape(handle):
    """512/2/32768/20: the factored sum over all 510 non-occupied bands
    against the per-term FP64 kernel on the same inputs, and a 3-band slice
    against the oracle."""
    say("factored canonical")
    inp = gppk.fill_inputs(2, 512, 32768, 20, nvband=2, seed=2024)
    sub = gppk.band_slice(inp, 0, 3)
    check(handle.sigma(sub, precision=FACT), orc.sigma(sub, want_achstemp=False))
    f = handle.sigma(inp, precision=FACT)
    d = handle.sigma(inp)
    for k in ("achtemp", "asxtemp"):
        assert rel_err(f[k], d[k]) <= 1e-12, (k, f[k], d[k])
    say(f"unocc kernel: factored {f['unocc_kernel_ms']:.3f} ms, "
        f"per-term {d['unocc_kernel_ms']:.3f} ms")


def test_factored_rejects_what_it_cannot_do(handle):
    inp3 = gppk.fill_inputs(3, 6, 256, 4, nvband=2, seed=5)
    with pytest.raises(gppk.GppError) as ei:
        handle.sigma(inp3, precision=FACT)
    assert ei.value.code == gppk.GPP_EUNSUPPORTED
    inp = gppk.fill_inputs(2, 6, 256, 4, nvband=2, seed=5)
    with pytest.raises(gppk.GppError) as ei:
        handle.autotune(inp, precision=FACT)
    assert ei.value.code == gppk.GPP_EUNSUPPORTED


@pytest.mark.parametrize("shape", [(10, 3, 700, 7), (70, 2, 2100, 10), (5, 0, 100, 3),
                                   (131, 4, 4500, 30), (9, 9, 300, 4)])
@pytest.mark.parametrize("sem", ["gppker", "strict", "bgw"])
def test_factored_acht_n1_loc(handle, sem, shape):
    """The per-band log in the factored mode (a second matrix product, over
    the G-vectors, for the non-occupied bands; the occupied bands' solver per
    band): against the oracle per band, several band tiles of 64 and
    G-vector chunks of 2048; it sums to achtemp[2], and asking for it leaves
    achtemp unchanged (bit for bit without occupied bands; with them their
    solver then runs band by band, another summation order)."""
    nb, nv, nc, npg = shape
    say("factored acht_n1_loc", shape, sem)
    inp = gppk.fill_inputs(2, nb, nc, npg, nvband=nv, seed=nb + nc)
    got = handle.sigma(inp, precision=FACT, semantics=sem, acht_n1_loc=True)
    ref = orc.sigma(inp, sem, want_acht_n1=True)
    check(got, ref)
    e = rel_err(got["acht_n1_loc"], ref["acht_n1_loc"])
    assert e <= FP64_TOL, e
    assert rel_err(np.sum(got["acht_n1_loc"]), got["achtemp"][2]) <= 1e-12
    plain = handle.sigma(inp, precision=FACT, semantics=sem)
    assert rel_err(got["achtemp"], plain["achtemp"]) <= 1e-13
    if nv == 0:
        assert np.array_equal(plain["achtemp"], got["achtemp"])


def test_factored_acht_n1_loc_skipped_rows_and_host_inputs(handle):
    """Skipped G' rows, and host inputs (chunked launches over G' groups)
    giving the device inputs' bits for the per-band log."""
    import torch
    say("factored acht_n1_loc host")
    inp = gppk.fill_inputs(2, 12, 8000, 4, nvband=2, seed=93)
    inv = inp["inv_igp_index"].copy()
    inv[5] = -1
    inv[700] = 8000
    q = dict(inp, inv_igp_index=inv)
    a = handle.sigma(q, precision=FACT, acht_n1_loc=True)
    dev = {k: (torch.from_numpy(np.ascontiguousarray(v)).cuda() if isinstance(v, np.ndarray) else v)
           for k, v in q.items()}
    torch.cuda.synchronize()
    b = handle.sigma(dev, precision=FACT, memory=gppk.MEM_DEVICE, acht_n1_loc=True)
    for k in ("achtemp", "asxtemp", "acht_n1_loc"):
        assert np.array_equal(a[k], b[k]), k
    ref = orc.sigma(q, want_acht_n1=True)
    assert rel_err(a["acht_n1_loc"], ref["acht_n1_loc"]) <= FP64_TOL


def test_factored_acht_n1_loc_canonical(handle):
    """512/2/32768/20: the factored per-band log against the per-term FP64
    kernel's, band by band."""
    say("factored acht_n1_loc canonical")
    inp = gppk.fill_inputs(2, 512, 32768, 20, nvband=2, seed=2024)
    f = handle.sigma(inp, precision=FACT, acht_n1_loc=True)
    d = handle.sigma(inp, acht_n1_loc=True)
    assert rel_err(f["acht_n1_loc"], d["acht_n1_loc"]) <= 1e-12
    say(f"with the per-band log: factored {f['unocc_kernel_ms']:.3f} ms, "
        f"per-term {d['unocc_kernel_ms']:.3f} ms")


@pytest.mark.parametrize("shape", [(1, 0, 1, 1), (2, 1, 1, 1), (1, 1, 5, 5), (3, 0, 2, 2),
                                   (6, 2, 65, 65), (4, 1, 63, 1), (65, 1, 64, 1)])
def test_factored_tiny_and_boundary_shapes(handle, shape):
    """One G-vector, one G' row, all bands occupied, tile boundaries (63, 64,
    65 G-vectors or G' rows; 64 non-occupied bands = one band tile) in every
    semantics, with the per-band log."""
    nb, nv, nc, npg = shape
    say("factored tiny", shape)
    inp = gppk.fill_inputs(2, nb, nc, npg, nvband=nv, seed=7 * nb + nc)
    for sem in ("gppker", "strict", "bgw"):
        got = handle.sigma(inp, precision=FACT, semantics=sem, acht_n1_loc=True)
        ref = orc.sigma(inp, sem, want_acht_n1=True)
        check(got, ref)
        assert rel_err(got["acht_n1_loc"], ref["acht_n1_loc"]) <= FP64_TOL


def test_factored_batch_runs_each_problem(handle):
    """A batch in the factored mode gives each problem's own factored
    result (the problems run one after another); band-resolved frequencies
    are refused as in a single call."""
    say("factored batch")
    inputs = [gppk.fill_inputs(2, nb, nc, 4, nvband=nv, seed=50 + i)
              for i, (nb, nc, nv) in enumerate([(6, 300, 2), (9, 700, 0), (3, 130, 3)])]
    res = handle.batch(inputs, precision=FACT)
    for d, got in zip(inputs, res):
        one = handle.sigma(d, precision=FACT)
        for k in ("achtemp", "asxtemp"):
            assert np.array_equal(got[k], one[k]), k
        assert got["achstemp"] == one["achstemp"]
        check(got, orc.sigma(d))
    with pytest.raises(gppk.GppError) as ei:
        handle.batch([inputs[0], gppk.fill_inputs(3, 4, 200, 4, nvband=1, seed=3)],
                     precision=FACT)
    assert ei.value.code == gppk.GPP_EUNSUPPORTED


def _wide_shapes(n, seed):
    """Random shapes with more G-vectors and G' rows than the ragged sweep
    (several 2048-wide chunks and 64-row groups), nvband up to
    number_bands."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nb = int(rng.integers(1, 48))
        nv = int(rng.integers(0, min(nb, 6) + 1))
        npg = int(rng.integers(6, 16))
        nc = int(rng.integers(64, 4500))
        out.append((nb, nv, nc, npg))
    return out


@pytest.mark.parametrize("shape", _wide_shapes(12, 77))
def test_factored_wide_sweep_with_log(handle, shape):
    nb, nv, nc, npg = shape
    say("factored wide sweep", shape)
    inp = gppk.fill_inputs(2, nb, nc, npg, nvband=nv, seed=3 * nb + nc)
    for sem in ("gppker", "bgw"):
        got = handle.sigma(inp, precision=FACT, semantics=sem, acht_n1_loc=True)
        ref = orc.sigma(inp, sem, want_acht_n1=True)
        check(got, ref)
        assert rel_err(got["acht_n1_loc"], ref["acht_n1_loc"]) <= FP64_TOL
