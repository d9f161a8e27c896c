"""Mixed precision (GPP_PREC_MIXED) on the GPU, through the C ABI.

The FP32 path replaces the FP64 non-occupied sum of gppKer.cpp:326-343;
near-resonant runs (and, under the bgw mask of
gppkernel.deslippe.09.30.2014.f90:364-366, runs near the mask's threshold)
are evaluated in FP64 through a per-wave work list.  Checked here:

  * the work list is exact: with a threshold above every |wdiff|^2 every run
    goes through it, and the result equals the oracle to the FP64 bar, in
    every semantics, tile, the per-band log and the batched launch;
  * the default path stays within its measured per-entry bound of the
    oracle on the canonical and band-resolved distributions;
  * a resonance window brings band-resolved inputs below 1e-8;
  * the bgw mask's threshold is decided in FP64 for terms next to it;
  * exact resonances give NaN in their own frequency only.
"""
import numpy as np
import pytest

from conftest import gppk, gpp_oracle as orc, rel_err
from test_oracle_cpu import masked_expected, masked_problem

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-11
ALL_FP64 = 1e30  # a threshold above every |wdiff|^2: every run in FP64

# Per-entry bounds of the default mixed path against the oracle on the
# shapes of this file, about 2.5x the worst measured (gppker 3.8e-7, bgw
# 8.3e-7: the mask drops the large near-resonant terms, so the FP32 error of
# the rest weighs more); the FP32 term error (a few 2^-24 per term,
# random-walking with the terms) sets this floor.
MIXED_BOUND = {"gppker": 1e-6, "strict": 1e-6, "bgw": 2e-6}


def say(*a):
    print(*a, flush=True)


def entry_err(got, ref, keys=("achtemp", "asxtemp")):
    return max(rel_err(got[k], ref[k]) for k in keys)


SHAPES = [  # (config, NB, NV, NC, nodes_per_group): ragged on purpose
    (2, 12, 2, 1000, 8), (3, 12, 2, 1000, 8), (3, 7, 3, 777, 7),
    (2, 5, 0, 130, 2), (3, 9, 9, 300, 3),
]


@pytest.mark.parametrize("sem", ["gppker", "strict", "bgw"])
@pytest.mark.parametrize("tile", [(8, 1), (4, 2), (16, 4), (32, 1)])
def test_work_list_is_exact(handle, sem, tile):
    k, w = tile
    for cfg, nb, nv, nc, npg in SHAPES:
        say("work list", sem, tile, (cfg, nb, nv, nc, npg))
        inp = gppk.fill_inputs(cfg, nb, nc, npg, nvband=nv, seed=nb + nc)
        ref = orc.sigma(inp, sem)
        got = handle.sigma(inp, precision="mixed", semantics=sem,
                           resonance_threshold=ALL_FP64, tile_ig=k,
                           waves_per_block=w)
        assert entry_err(got, ref) <= FP64_TOL, (got, ref)
        nu = (nb - min(nv, nb)) * (nc // npg) * nc * 3
        # every non-occupied term went through the list (inert slots of a
        # run past ncouls are counted with their run)
        assert got["fp64_fallback_terms"] >= nu


@pytest.mark.parametrize("sem", ["gppker", "bgw"])
def test_work_list_per_band_log(handle, sem):
    inp = gppk.fill_inputs(3, 12, 1000, 8, nvband=2, seed=5)
    ref = orc.sigma(inp, sem, want_acht_n1=True)
    got = handle.sigma(inp, precision="mixed", semantics=sem,
                       resonance_threshold=ALL_FP64, acht_n1_loc=True)
    assert rel_err(got["acht_n1_loc"], ref["acht_n1_loc"]) <= FP64_TOL
    got = handle.sigma(inp, precision="mixed", semantics=sem, acht_n1_loc=True)
    e = rel_err(got["acht_n1_loc"], ref["acht_n1_loc"])
    say(f"per-band log, default threshold: {e:.2e}")
    assert e <= MIXED_BOUND[sem]
    assert rel_err(np.sum(got["acht_n1_loc"]), got["achtemp"][2]) <= 1e-12


@pytest.mark.parametrize("sem", ["gppker", "bgw"])
def test_work_list_in_batches(handle, sem):
    inputs = [gppk.fill_inputs(2 + i % 2, nb, nc, 4, nvband=nv, seed=60 + i)
              for i, (nb, nc, nv) in enumerate([(5, 300, 1), (9, 130, 0),
                                                (3, 777, 2), (12, 64, 3)])]
    refs = [orc.sigma(d, sem) for d in inputs]
    for rep in range(2):
        res = handle.batch(inputs, precision="mixed", semantics=sem,
                           resonance_threshold=ALL_FP64)
        for got, ref in zip(res, refs):
            assert entry_err(got, ref) <= FP64_TOL
    res = handle.batch(inputs, precision="mixed", semantics=sem)
    for got, ref in zip(res, refs):
        assert entry_err(got, ref) <= MIXED_BOUND[sem]


@pytest.mark.parametrize("kahan", [False, True])
@pytest.mark.parametrize("sem", ["gppker", "bgw"])
def test_default_path_bound(handle, sem, kahan):
    flags = gppk.FLAG_MIXED_KAHAN if kahan else 0
    worst = 0.0
    for cfg, nb, nv, nc, npg in SHAPES + [(2, 64, 2, 4096, 16), (3, 64, 2, 4096, 16)]:
        inp = gppk.fill_inputs(cfg, nb, nc, npg, nvband=nv, seed=3 * nb + nc)
        ref = orc.sigma(inp, sem)
        got = handle.sigma(inp, precision="mixed", semantics=sem, flags=flags)
        e = entry_err(got, ref)
        say(f"mixed {sem} kahan={kahan} config {cfg} {nb}/{nv}/{nc}/{npg}: {e:.2e}")
        worst = max(worst, e)
        # occupied bands run in FP64
        assert rel_err(got["asxtemp"], ref["asxtemp"]) <= FP64_TOL or nv == 0
    assert worst <= MIXED_BOUND[sem], worst


def test_resonance_window_band_resolved(handle):
    """Band-resolved energies with 5% near-resonant wtilde, 256 bands: the
    near-resonant terms dominate the sum, and evaluating the runs that hold
    a |wdiff|^2 < 4e-3 in FP64 (about 5% of the terms) brings every entry
    below 1e-8 of the oracle (measured at full size: 5.5e-9 against 1.9e-8
    at the default threshold)."""
    inp = gppk.fill_inputs(3, 256, 4096, 8, nvband=2, seed=2024)
    ref = orc.sigma(inp)
    base = handle.sigma(inp, precision="mixed")
    win = handle.sigma(inp, precision="mixed", resonance_threshold=4e-3)
    e0, e1 = entry_err(base, ref), entry_err(win, ref)
    frac = win["fp64_fallback_terms"] / win["terms"]
    say(f"band-resolved 256/2/4096/8: default {e0:.2e}, window 4e-3 {e1:.2e} "
        f"({frac:.3f} of the terms in FP64)")
    assert e1 <= 1e-8
    assert e1 < e0
    assert 0.0 < frac < 0.15


def test_bgw_threshold_decided_in_fp64(handle):
    """wtilde = 4.5 + 1e-12 against wx = 4 and 5: |wdiff|^2 = 0.25 + 1e-12
    (kept by the f90 mask, > 0.25) and 0.25 - 1e-12 (dropped) in FP64, while
    both FP32 values are exactly 0.25.  The run is evaluated in FP64 and the
    result matches the oracle and the closed form; wtilde = 4.5 drops both."""
    for wp, dropped_terms in ((4.5 + 1e-12, 3), (4.5, 6)):
        placed = [(2, 11, complex(wp))]
        inp = masked_problem(placed)
        ref = orc.sigma(inp, "bgw")
        want, dropped = masked_expected(inp, placed, "bgw")
        assert dropped == dropped_terms  # per band: 1 or 2 of 3 frequencies
        got = handle.sigma(inp, precision="mixed", semantics="bgw")
        assert rel_err(got["achtemp"], ref["achtemp"]) <= MIXED_BOUND["bgw"]
        assert rel_err(got["achtemp"], want) <= MIXED_BOUND["bgw"]
        assert got["fp64_fallback_terms"] > 0
    # the kept term moves achtemp[1] by far more than the bound
    inp = masked_problem([(2, 11, 4.5 + 1e-12 + 0j)])
    kept_val = handle.sigma(inp, precision="mixed", semantics="bgw")["achtemp"][1]
    inp = masked_problem([(2, 11, 4.5 + 0j)])
    dropped_val = handle.sigma(inp, precision="mixed", semantics="bgw")["achtemp"][1]
    assert abs(kept_val - dropped_val) > 1e3 * MIXED_BOUND["bgw"] * abs(kept_val)


def test_exact_resonance_nan_in_its_frequency_only(handle):
    inp = gppk.fill_inputs(2, 3, 64, 4, nvband=0, seed=3)
    w = inp["wtilde_array"].copy()
    w[0, 5] = 4.0 + 0.0j  # exactly wx[1]
    got = handle.sigma(dict(inp, wtilde_array=w), precision="mixed")
    assert np.isnan(got["achtemp"][1])
    assert np.all(np.isfinite(got["achtemp"][[0, 2]]))
    assert got["fp64_fallback_terms"] > 0


@pytest.mark.parametrize("sem", ["gppker", "bgw"])
def test_nan_frequency_in_its_entry_only(handle, sem):
    """A NaN frequency of one band (band-resolved wx): the FP32 path forms
    the three frequencies' reciprocals together, so the run goes to FP64 and
    the NaN stays where the oracle has it (its own entry under gppker; the
    bgw mask drops the term, as the f90 comparisons do); the other entries
    match the oracle."""
    inp = gppk.fill_inputs(3, 6, 256, 4, nvband=1, seed=9)
    wx = inp["wx_array"].copy()
    wx[3, 1] = np.nan
    inp = dict(inp, wx_array=wx)
    ref = orc.sigma(inp, sem)
    got = handle.sigma(inp, precision="mixed", semantics=sem)
    nan = np.isnan(ref["achtemp"])
    assert nan[1] == (sem == "gppker") and not nan[0] and not nan[2]
    assert np.array_equal(np.isnan(got["achtemp"]), nan)
    keep = ~nan
    assert rel_err(got["achtemp"][keep], ref["achtemp"][keep]) <= MIXED_BOUND[sem]
    assert got["fp64_fallback_terms"] > 0
    # the FP64 kernel's per-term fallback treats the band the same way
    f64 = handle.sigma(inp, precision="fp64", semantics=sem)
    assert np.array_equal(np.isnan(f64["achtemp"]), nan)
    assert rel_err(f64["achtemp"][keep], ref["achtemp"][keep]) <= FP64_TOL


@pytest.mark.parametrize("sem", ["gppker", "bgw"])
def test_nan_wtilde(handle, sem):
    """A NaN wtilde: its lane's runs go to FP64, where (as in the oracle)
    the bgw mask drops its terms and gppker keeps them (NaN in every
    entry); the FP64 kernel's fallback agrees."""
    inp = gppk.fill_inputs(3, 6, 256, 4, nvband=1, seed=10)
    w = inp["wtilde_array"].copy()
    w[5, 17] = complex(np.nan, 0.02)
    inp = dict(inp, wtilde_array=w)
    ref = orc.sigma(inp, sem)
    nan = np.isnan(ref["achtemp"])
    assert nan.all() == (sem == "gppker") and (sem == "gppker" or not nan.any())
    for prec, tol in (("mixed", MIXED_BOUND[sem]), ("fp64", FP64_TOL)):
        got = handle.sigma(inp, precision=prec, semantics=sem)
        assert np.array_equal(np.isnan(got["achtemp"]), nan), prec
        if not nan.any():
            assert rel_err(got["achtemp"], ref["achtemp"]) <= tol, prec


def test_large_values_stay_within_the_bound(handle):
    """wtilde components and frequencies of 1e5 and more: FP32 terms of
    |wdiff|^2 up to ~1e13 stay finite and within the mixed bound."""
    inp = gppk.fill_inputs(3, 7, 300, 3, nvband=2, seed=11)
    w = inp["wtilde_array"].copy()
    w[1, 7] = 3e5 + 0.01j
    w[2, 8] = 1.5 + 2e6j
    wx = inp["wx_array"].copy()
    wx[4, :] = [4e5, 4e5 + 1.0, 4e5 + 2.0]
    inp = dict(inp, wtilde_array=w, wx_array=wx)
    ref = orc.sigma(inp)
    got = handle.sigma(inp, precision="mixed")
    assert np.all(np.isfinite(got["achtemp"]))
    assert entry_err(got, ref) <= MIXED_BOUND["gppker"]
