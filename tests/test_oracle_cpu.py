"""CPU tests: the oracle pinned against the reference and against the
mathematics; the C-ABI library loads and exports every declared symbol.

Pins used (each fails on a plausible oracle mistake):
  * ref_constfill.json   achtemp printed by the reference binary
                         (gppKerOpenMPFor_xl.cpp) for its constant fills;
  * ref_occ_kat_*.json   the reference's own flagOCC_solver on random inputs,
                         both I_eps index variants (BGW-Kernels/gppKer.cpp:155
                         and gppKer.cpp:139);
  * closed forms for the constant fills, both semantics;
  * invariants: linearity in vcoul and aqsn, conjugate-linearity in aqsm,
    band additivity, ig permutation, nvband = 0.
"""
import ctypes
import json
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, gppk, gpp_oracle as orc, rel_err

C = 0.5 + 0.5j


def closed_form(nb, nv, nc, npg, sem="gppker"):
    """Constant fills c = 0.5+0.5i, vcoul = 1 (gppKer.cpp:241-264):
    every term of every band equals c*conj(c)*c*delw, and the occupied
    terms' ssx are c^3/wx^2 (gppker) or c^3 conj(wx^2-c^2)/|wx^2-c^2|^2
    (bgw), times |c|^2 from matngmatmgp."""
    ng = nc // npg
    ach, asx = [], []
    for wx in (3.0, 4.0, 5.0):
        delw = C * np.conj(wx - C) / abs(wx - C) ** 2
        ach.append(nb * ng * nc * 0.25 * C * delw)
        if sem == "bgw":
            cden = wx * wx - C * C
            ssx = C ** 3 * np.conj(cden) / abs(cden) ** 2
            asx.append(-min(nv, nb) * ng * nc * 0.5 * ssx)
        else:
            asx.append(min(nv, nb) * ng * nc * 0.5 * C ** 3 / wx ** 2)
    return np.array(ach), np.array(asx)


def small_random(nb=6, nv=2, nc=48, npg=4, seed=3, config=2):
    return gppk.fill_inputs(config, nb, nc, npg, nvband=nv, seed=seed)


# --------------------------------------------------------------------------
# Library boundary


def test_library_exports_every_declared_symbol():
    lib = gppk.load_library()
    declared = set()
    for name in ("gpp_sigma.h", "ff_sigma.h"):
        header = (ROOT / "include" / name).read_text()
        declared |= set(re.findall(
            r"^\s*(?:GPP_API\s+)?(?:int|void|const char \*)\s*((?:gpp|ff)_\w+)\(",
            header, re.M))
    assert declared == set(gppk.EXPORTS), declared ^ set(gppk.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name
    assert b"gfx950" in lib.gpp_version()
    # and nothing else: internal C and C++ symbols stay hidden
    import shutil, subprocess
    if shutil.which("nm"):
        out = subprocess.run(["nm", "-D", "--defined-only", str(gppk.LIB_PATH)],
                             capture_output=True, text=True, check=True).stdout
        exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
        exported = {e for e in exported if not e.startswith("__hip_cuid_")}
        assert exported == declared, exported ^ declared


def test_fill_inputs_is_deterministic_and_shaped():
    a = small_random(seed=5)
    b = small_random(seed=5)
    c = small_random(seed=6)
    for k in ("aqsmtemp", "aqsntemp", "I_eps_array", "wtilde_array", "vcoul"):
        assert np.array_equal(a[k], b[k])
    assert not np.array_equal(a["aqsmtemp"], c["aqsmtemp"])
    assert a["aqsmtemp"].shape == (6, 48) and a["I_eps_array"].shape == (12, 48)
    assert np.all(a["inv_igp_index"] == 4 * np.arange(12))
    w = a["wtilde_array"]
    assert np.all((w.real >= 0.5) & (w.real <= 3.0) & (np.abs(w.imag) <= 0.05))
    e = a["I_eps_array"]
    assert np.all((e.real >= 0) & (e.real < 1) & (np.abs(e.imag) <= 0.1))
    ig = np.arange(48)
    assert np.allclose(a["vcoul"], 1.0 / (0.01 + (ig / 48) ** (2 / 3)))
    assert np.array_equal(a["wx_array"], [3.0, 4.0, 5.0])


def test_fill_inputs_band_resolved():
    d = small_random(nb=40, nc=256, npg=4, config=3)
    wx = d["wx_array"]
    assert wx.shape == (40, 3)
    e = 4.0 - wx[:, 1]
    assert np.all(np.diff(e) >= 0) and np.all((e >= -20) & (e < 20))
    assert np.allclose(wx[:, 0], wx[:, 1] - 1) and np.allclose(wx[:, 2], wx[:, 1] + 1)
    # about 5% of wtilde moved within 1e-4 of some wx
    near = np.min(np.abs(d["wtilde_array"].real[..., None] - wx.ravel()), axis=-1) <= 1e-4
    frac = near.mean()
    assert 0.03 < frac < 0.07, frac


def test_cli_usage_and_constant_fill_output():
    import subprocess
    exe = gppk.CLI_PATH
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0
    assert "The correct form of input is" in out.stdout


# --------------------------------------------------------------------------
# Oracle vs the reference binary and closed forms


@pytest.mark.parametrize("rec", json.loads((GOLDEN / "ref_constfill.json").read_text()),
                         ids=lambda r: "{number_bands}-{nvband}-{ncouls}-{nodes_per_group}".format(**r))
def test_oracle_matches_reference_binary_constfill(rec):
    nb, nv, nc, npg = rec["number_bands"], rec["nvband"], rec["ncouls"], rec["nodes_per_group"]
    inp = gppk.fill_inputs(1, nb, nc, npg, nvband=nv)
    # The reference's own strided index table (gppKer.cpp:259-260)
    ng = nc // npg
    assert np.array_equal(inp["inv_igp_index"], (np.arange(ng) + 1) * nc // ng)
    ref = np.array([complex(*z) for z in rec["achtemp"]])
    for sem in ("gppker", "strict", "bgw"):
        got = orc.sigma(inp, sem)["achtemp"]
        # printed at 15 significant digits after a plain (uncompensated) sum
        assert rel_err(got, ref) < 5e-12, (sem, got, ref)


@pytest.mark.parametrize("shape", [(64, 2, 2048, 8), (8, 2, 64, 4), (5, 5, 40, 8),
                                   (12, 0, 100, 7), (3, 7, 30, 3)])
@pytest.mark.parametrize("sem", ["gppker", "bgw"])
def test_oracle_closed_form(shape, sem):
    nb, nv, nc, npg = shape
    inp = gppk.fill_inputs(1, nb, nc, npg, nvband=nv)
    got = orc.sigma(inp, sem)
    ach, asx = closed_form(nb, nv, nc, npg, sem)
    assert rel_err(got["achtemp"], ach) < 1e-14
    if min(nv, nb) == 0:
        assert np.all(got["asxtemp"] == 0)
    else:
        assert rel_err(got["asxtemp"], asx) < 1e-14


def test_oracle_closed_form_values_64_2_2048_8():
    """The mini-app fills at 64/2/2048/8: achtemp[iw] = NB (NC/4) c delw(wx)
    sum_igp vcoul and asxtemp[iw] = NV NG NC c^3 / (2 wx^2), c = 0.5+0.5i,
    wx = 3, 4, 5, evaluated by hand."""
    inp = gppk.fill_inputs(1, 64, 2048, 8, nvband=2)
    got = orc.sigma(inp)
    want_ach = [-322638.769230769 + 1613193.84615385j, -167772.16 + 1174405.12j,
                -102300.097560976 + 920700.87804878j]
    want_asx = [-14563.5555555556 + 14563.5555555556j, -8192 + 8192j,
                -5242.88 + 5242.88j]
    assert rel_err(got["achtemp"], want_ach) < 1e-12
    assert rel_err(got["asxtemp"], want_asx) < 1e-12


# --------------------------------------------------------------------------
# Oracle vs the reference's own occupied solver on random inputs


def _kat(kind):
    d = json.loads((GOLDEN / f"ref_occ_kat_{kind}.json").read_text())
    c = lambda a: np.array(a)[..., 0] + 1j * np.array(a)[..., 1]
    nb, nc, ng = d["number_bands"], d["ncouls"], d["ngpown"]
    inp = {"aqsmtemp": c(d["aqsmtemp"]), "aqsntemp": c(d["aqsntemp"]),
           "I_eps_array": c(d["I_eps_array"]), "wtilde_array": c(d["wtilde_array"]),
           "vcoul": np.ones(nc), "inv_igp_index": np.array(d["igp"], np.int32),
           "indinv": np.arange(nc, dtype=np.int32),
           "wx_array": np.array(d["wx_array"]), "nvband": nb}
    return d, inp


@pytest.mark.parametrize("kind", ["branches", "masktrue"])
@pytest.mark.parametrize("variant,sem", [("bgwk", "gppker"), ("xl", "strict")])
def test_oracle_occ_rows_match_reference_solver(kind, variant, sem):
    d, inp = _kat(kind)
    worst = 0.0
    scale = max(abs(complex(r[3], r[4])) for r in d["rows_" + variant])
    for n1, my_igp, iw, ssr, ssi, scr, sci in d["rows_" + variant]:
        wxt = d["wx_array"][n1][iw]
        ssxt, scht = orc.occ_sums(inp, sem, n1, my_igp, d["igp"][my_igp], wxt)
        worst = max(worst, abs(ssxt - complex(ssr, ssi)) / scale,
                    abs(scht - complex(scr, sci)) / scale)
    assert worst < 1e-13, worst


def test_kat_variants_really_differ():
    # The two index variants disagree for my_igp > 0: the fixture pins both.
    d, _ = _kat("branches")
    a = np.array(d["rows_bgwk"])[:, 5:]
    b = np.array(d["rows_xl"])[:, 5:]
    assert not np.allclose(a, b)


def test_oracle_unocc_sum_matches_reference_solver_when_mask_true():
    """With |wx - wtilde|^2 > 0.25 everywhere, the reference solver's scht is
    sum_ig conj(aqsm)*aqsn*delw*I_eps: the non-occupied term of
    gppKer.cpp:337.  The oracle's non-occupied path, run on the same inputs
    with every band unoccupied, must reproduce 0.5*sum vcoul*scht."""
    d, inp = _kat("masktrue")
    rows = np.array(d["rows_bgwk"])
    want = np.zeros(3, complex)
    for n1, my_igp, iw, _, _, scr, sci in rows:
        want[int(iw)] += 0.5 * complex(scr, sci)  # vcoul = 1
    got = orc.sigma(inp, "gppker", flags=orc.ALL_BANDS_UNOCC)["achtemp"]
    assert rel_err(got, want) < 1e-13
    # and the occupied path, every band occupied
    got2 = orc.sigma(inp, "gppker")["achtemp"]
    assert rel_err(got2, want) < 1e-13


def test_oracle_asxtemp_from_reference_solver():
    d, inp = _kat("branches")
    for variant, sem in (("bgwk", "gppker"), ("xl", "strict")):
        want = np.zeros(3, complex)
        ach = np.zeros(3, complex)
        for n1, my_igp, iw, ssr, ssi, scr, sci in d["rows_" + variant]:
            want[int(iw)] += complex(ssr, ssi)          # occ * vcoul = 1
            ach[int(iw)] += 0.5 * complex(scr, sci)
        got = orc.sigma(inp, sem)
        assert rel_err(got["asxtemp"], want) < 1e-13
        assert rel_err(got["achtemp"], ach) < 1e-13


# --------------------------------------------------------------------------
# Invariants (random inputs)


def test_linearity_in_vcoul_and_aqsn():
    inp = small_random(nb=5, nv=2, nc=40)
    base = orc.sigma(inp)
    v2 = dict(inp, vcoul=inp["vcoul"] * 3.0)
    r = orc.sigma(v2)
    assert rel_err(r["achtemp"], 3 * base["achtemp"]) < 1e-13
    assert rel_err(r["asxtemp"], 3 * base["asxtemp"]) < 1e-13
    s = 0.3 - 1.7j
    r = orc.sigma(dict(inp, aqsntemp=inp["aqsntemp"] * s))
    assert rel_err(r["achtemp"], s * base["achtemp"]) < 1e-13
    assert rel_err(r["asxtemp"], s * base["asxtemp"]) < 1e-13
    # conjugate-linear in aqsm
    r = orc.sigma(dict(inp, aqsmtemp=inp["aqsmtemp"] * s))
    assert rel_err(r["achtemp"], np.conj(s) * base["achtemp"]) < 1e-13


def test_band_additivity():
    inp = small_random(nb=7, nv=3, nc=40, config=3)
    full = orc.sigma(inp)
    parts = [orc.sigma(gppk.band_slice(inp, a, b)) for a, b in ((0, 2), (2, 5), (5, 7))]
    for k in ("achtemp", "asxtemp"):
        assert rel_err(sum(p[k] for p in parts), full[k]) < 1e-13


def test_ig_permutation_invariance():
    inp = small_random(nb=4, nv=2, nc=32)
    rng = np.random.default_rng(0)
    perm = rng.permutation(32)
    inv = np.argsort(perm)
    q = dict(inp)
    for k in ("aqsmtemp", "aqsntemp", "I_eps_array", "wtilde_array"):
        q[k] = np.ascontiguousarray(inp[k][:, perm])
    q["vcoul"] = inp["vcoul"][perm]
    # igp moves with the permutation: new position of old igp
    q["indinv"] = inv[inp["indinv"]].astype(np.int32)
    a, b = orc.sigma(inp), orc.sigma(q)
    for k in ("achtemp", "asxtemp"):
        assert rel_err(b[k], a[k]) < 1e-13


def test_nvband_zero_and_all():
    inp = small_random(nb=4, nv=0, nc=32)
    r = orc.sigma(inp)
    assert np.all(r["asxtemp"] == 0)
    un = orc.sigma(inp, flags=orc.ALL_BANDS_UNOCC)
    assert rel_err(r["achtemp"], un["achtemp"]) == 0.0
    r2 = orc.sigma(dict(inp, nvband=99))
    assert np.all(r2["asxtemp"] != 0)


def test_skipped_rows_and_index_edge():
    inp = small_random(nb=3, nv=1, nc=32, npg=4)
    inv = inp["inv_igp_index"].copy()
    inv[1] = -5       # out of table -> row skipped
    inv[2] = 32       # == ncouls -> igp = ncouls - 1 (gppKer.cpp:289-290)
    q = dict(inp, inv_igp_index=inv)
    igp = orc.resolve_igp(q)
    assert igp[1] == -1 and igp[2] == 31
    # skipping row 1 equals zeroing its weights through vcoul is not possible
    # (vcoul is per igp); instead compare with the row removed.
    keep = [i for i in range(len(inv)) if i != 1]
    r = dict(q, inv_igp_index=inv[keep], I_eps_array=q["I_eps_array"][keep],
             wtilde_array=q["wtilde_array"][keep], ngpown=len(keep))
    a, b = orc.sigma(q), orc.sigma(r)
    for k in ("achtemp", "asxtemp"):
        assert rel_err(a[k], b[k]) < 1e-14


def test_exact_resonance_gives_nan_in_that_frequency_only():
    inp = small_random(nb=3, nv=0, nc=32)
    w = inp["wtilde_array"].copy()
    w[0, 5] = 4.0 + 0.0j  # exactly wx[1]
    r = orc.sigma(dict(inp, wtilde_array=w))
    assert np.isnan(r["achtemp"][1])
    assert np.all(np.isfinite(r["achtemp"][[0, 2]]))


def test_bgw_mask_drops_resonant_terms():
    inp = small_random(nb=3, nv=0, nc=32)
    w = inp["wtilde_array"].copy()
    w[0, 5] = 4.0 + 0.0j
    r = orc.sigma(dict(inp, wtilde_array=w), "bgw")
    assert np.all(np.isfinite(r["achtemp"]))


# --------------------------------------------------------------------------
# Closed forms for what the constant fills cannot reach: the bgw cutoff
# (f90:321, only for wx < 0) and the static remainder achstemp.

CUT_WX = np.array([[-3.48, 3.48, -5.0]])


def cutoff_problem():
    """One occupied band, constant fills except wtilde = 4 (real) and
    band-resolved wx = (-3.48, 3.48, -5): every term takes the solver's
    first branch (|wdiff|^2 = (wx - 4)^2 > 0.25, |delw|^2 < 2.5e5) with
    Omega2 = 16 I_eps, so per term
      gppker: ssx = Omega2 / wx^2 (its cutoff tests abs(wx) < 0: never),
      bgw:    ssx = Omega2 conj(cden) / |cden|^2 = 16 c / (wx^2 - 16),
              zeroed when wx < 0 and |ssx| > 4 |I_eps|:
    |ssx| = 4.11 |c| at wx = -3.48 (cut), the same at +3.48 (kept: wx > 0),
    1.78 |c| at -5 (kept)."""
    inp = gppk.fill_inputs(1, 1, 64, 8, nvband=1)
    return dict(inp, wtilde_array=np.full_like(inp["wtilde_array"], 4.0),
                wx_array=CUT_WX.copy())


def cutoff_expected(sem, inp):
    ng, nc = inp["I_eps_array"].shape
    wx = CUT_WX[0]
    k = ng * nc * abs(C) ** 2  # sum of matngmatmgp, occ = vcoul = 1
    if sem == "bgw":
        ssx = 16 * C / (wx * wx - 16)
        ssx = np.where((wx < 0) & (np.abs(ssx) > 4 * abs(C)), 0.0, ssx)
        return -k * ssx
    return k * 16 * C / wx ** 2


@pytest.mark.parametrize("sem", ["gppker", "bgw"])
def test_oracle_occupied_cutoff_closed_form(sem):
    inp = cutoff_problem()
    got = orc.sigma(inp, sem)["asxtemp"]
    want = cutoff_expected(sem, inp)
    if sem == "bgw":
        assert got[0] == 0 and want[0] == 0
    assert rel_err(got, want) <= 1e-13, (got, want)


def achstemp_problem():
    """Random aqsm, aqsn, vcoul; I_eps zero except three entries:
    (0, igp(0)) = e1 on the igp column, (2, 5) = e2 off it (5 != igp(2)),
    and (3, igp(3)) = 1e-8 (below gppKer.cpp's |schs| > 1e-6 test)."""
    inp = small_random(nb=5, nv=2, nc=48, npg=4, seed=8)
    ieps = np.zeros_like(inp["I_eps_array"])
    igp = [int(inp["indinv"][inp["inv_igp_index"][m]]) for m in range(4)]
    assert igp[2] != 5
    ieps[0, igp[0]] = 0.7 - 0.2j
    ieps[2, 5] = -0.3 + 0.9j
    ieps[3, igp[3]] = 1e-8
    return dict(inp, I_eps_array=ieps), igp


def achstemp_expected(sem, inp, igp):
    """gppKer.cpp:93-97: sum over bands and rows of aqsn[n1][igp]
    conj(aqsm[n1][igp]) I_eps[my_igp][igp] vcoul[igp] / 2 when
    |I_eps[my_igp][igp]| > 1e-6; f90:214-224: minus the sum over ig of
    aqsn[n1][ig] I_eps[my_igp][ig] conj(aqsm[n1][igp]) vcoul[igp] / 2."""
    am, an, ie, vc = (inp["aqsmtemp"], inp["aqsntemp"], inp["I_eps_array"],
                      inp["vcoul"])
    total = 0j
    for m, ig in ((0, igp[0]), (2, 5), (3, igp[3])):
        g = igp[m]
        e = ie[m, ig]
        for n1 in range(am.shape[0]):
            if sem == "bgw":
                total -= an[n1, ig] * e * np.conj(am[n1, g]) * vc[g] * 0.5
            elif ig == g and abs(e) > 1e-6:
                total += an[n1, g] * np.conj(am[n1, g]) * e * vc[g] * 0.5
    return total


@pytest.mark.parametrize("sem", ["gppker", "bgw"])
def test_oracle_achstemp_single_entries(sem):
    """Pins which I_eps entries enter the static remainder (the igp column
    only, above 1e-6, in gppKer.cpp; every ig in the F90), the aqsn/aqsm
    indices and the signs."""
    inp, igp = achstemp_problem()
    got = orc.sigma(inp, sem)["achstemp"]
    want = achstemp_expected(sem, inp, igp)
    assert abs(got - want) <= 1e-14 * abs(want), (got, want)


@pytest.mark.parametrize("sem", ["gppker", "bgw"])
def test_oracle_acht_n1_loc_is_each_bands_achtemp2(sem):
    """acht_n1_loc[n1] (gppKer.cpp:357) is band n1's share of achtemp[2]:
    the achtemp[2] of the one-band sub-problem (band additivity), for
    occupied and unoccupied bands alike."""
    inp = small_random(nb=5, nv=2, nc=48, npg=4, seed=11, config=3)
    full = orc.sigma(inp, sem, want_acht_n1=True)["acht_n1_loc"]
    for n1 in range(5):
        one = orc.sigma(gppk.band_slice(inp, n1, n1 + 1), sem)["achtemp"][2]
        assert abs(full[n1] - one) <= 1e-13 * abs(one), (n1, full[n1], one)


def test_batch_result_view_matches_struct_fields():
    """The bulk numpy view of a gpp_result array used for batch results
    reads the same values as the field-by-field ctypes conversion."""
    import ctypes as C
    res = (gppk.Result * 3)()
    for i in range(3):
        for iw in range(3):
            res[i].achtemp[iw].re, res[i].achtemp[iw].im = i + iw, -iw - 0.5
            res[i].asxtemp[iw].re, res[i].asxtemp[iw].im = 2.0 * iw, i * 0.25
        res[i].achstemp.re, res[i].achstemp.im = 1.0 + i, -3.0
        res[i].unocc_kernel_ms, res[i].occ_kernel_ms = 1.5 * i, 0.5
        res[i].total_ms, res[i].terms = 2.5, 1e9 + i
        res[i].fp64_fallback_terms = 7 * i
    fast = gppk._batch_results(res, 3)
    for i in range(3):
        ref = gppk.Handle._result(res[i])
        assert set(fast[i]) == set(ref)
        for k, v in ref.items():
            assert np.array_equal(np.asarray(fast[i][k]), np.asarray(v)), (i, k)
    assert gppk._RESULT_DTYPE.itemsize == C.sizeof(gppk.Result)


# --------------------------------------------------------------------------
# The frequency clamp (gppKer.cpp:279: `if (abs(wx) < to1) wx = to1`), the
# linear Coulomb factor, and the bgw mask on a known subset of terms, by
# closed forms of the constant fills.

TO1 = 1e-6


def clamp(w):
    """gppKer.cpp:279 as written: any |wx| below to1 becomes +to1."""
    return TO1 if abs(w) < TO1 else w


def closed_form_rows(wx_rows, nv, ng, nc, vsum, sem="gppker"):
    """Constant fills with band-resolved frequencies wx_rows[n1][iw] (before
    the clamp) and sum_my_igp vcoul[igp] = vsum: every band adds
    vsum nc c delw(wx)/4 to achtemp (the occupied ones through the solver's
    first branch, which every wx here takes: |wx - c|^2 > 0.25 and
    |delw|^2 < 2), each occupied band vsum nc ssx(wx)/2 to asxtemp (see
    closed_form)."""
    ach = np.zeros(3, complex)
    asx = np.zeros(3, complex)
    for n1, row in enumerate(wx_rows):
        for iw in range(3):
            wx = clamp(row[iw])
            delw = C * np.conj(wx - C) / abs(wx - C) ** 2
            ach[iw] += vsum * nc * 0.25 * C * delw
            if n1 >= nv:
                continue
            if sem == "bgw":
                cden = wx * wx - C * C
                ssx = C ** 3 * np.conj(cden) / abs(cden) ** 2
                cut = wx < 0 and abs(ssx) > 4 * abs(C)
                asx[iw] -= 0.0 if cut else vsum * nc * 0.5 * ssx
            else:
                asx[iw] += vsum * nc * 0.5 * C ** 3 / wx ** 2
    return ach, asx


CLAMP_ROWS = [
    (0.0, -5e-7, 5e-7),        # all three clamped to +1e-6
    (1e-6, -1e-6, 2e-6),       # on and just past the threshold: unchanged
    (-3e-7, 3.0, 9.99e-7),     # clamped, far, clamped
    (3.0, 4.0, 5.0),           # the mini-app's grid
]


@pytest.mark.parametrize("sem", ["gppker", "bgw"])
@pytest.mark.parametrize("nv", [0, 2])
def test_oracle_clamps_small_frequencies(sem, nv):
    """Band-resolved rows with |wx| below, at and above 1e-6 in both
    signs, occupied and non-occupied bands.  Planting a missing clamp, a
    sign-keeping clamp (-5e-7 -> -1e-6) or `<=` for `<` (-1e-6 -> +1e-6)
    into the closed form moves achtemp or asxtemp by >= 1e-6 relative."""
    nb, nc, npg = len(CLAMP_ROWS), 64, 8
    inp = gppk.fill_inputs(1, nb, nc, npg, nvband=nv)
    inp["wx_array"] = np.array(CLAMP_ROWS, np.float64)
    got = orc.sigma(inp, sem)
    ach, asx = closed_form_rows(CLAMP_ROWS, nv, nc // npg, nc, nc // npg, sem)
    assert rel_err(got["achtemp"], ach) < 1e-12, (got["achtemp"], ach)
    if nv:
        assert rel_err(got["asxtemp"], asx) < 1e-12, (got["asxtemp"], asx)
    # the planted variants really are that far away
    def planted(fn):
        global clamp
        keep = clamp
        clamp = fn
        try:
            return closed_form_rows(CLAMP_ROWS, nv, nc // npg, nc, nc // npg, sem)
        finally:
            clamp = keep
    for fn in (lambda w: w,
               lambda w: np.copysign(TO1, w) if abs(w) < TO1 else w,
               lambda w: TO1 if abs(w) <= TO1 else w):
        try:
            a2, s2 = planted(fn)
        except ZeroDivisionError:  # unclamped wx = 0: ssx = c^3/0
            continue
        moved = rel_err(a2, ach) if not nv else max(rel_err(a2, ach), rel_err(s2, asx))
        assert moved > 1e-7, moved


@pytest.mark.parametrize("sem", ["gppker", "bgw"])
def test_oracle_clamps_shared_frequencies(sem):
    wx = np.array([0.0, -5e-7, 4.0])
    inp = gppk.fill_inputs(1, 5, 64, 8, nvband=2)
    inp["wx_array"] = wx.copy()
    got = orc.sigma(inp, sem)
    ach, asx = closed_form_rows([tuple(wx)] * 5, 2, 8, 64, 8, sem)
    assert rel_err(got["achtemp"], ach) < 1e-12
    assert rel_err(got["asxtemp"], asx) < 1e-12


def linear_vcoul_sum(nc, ng):
    """sum over my_igp of vcoul[igp] = igp/nc for the mini-app's strided
    G' columns inv_igp_index[k] = (k+1) nc/ng, the last one (== nc) mapped to
    nc-1 (gppKer.cpp:259-260,287-290)."""
    s = 0.0
    for k in range(ng):
        igp = (k + 1) * nc // ng
        if igp == nc:
            igp = nc - 1
        s += igp / nc
    return s


def test_fill_inputs_linear_vcoul():
    for cfg in (1, 2, 3):
        a = gppk.fill_inputs(cfg, 4, 96, 8, nvband=1, seed=9)
        b = gppk.fill_inputs(cfg, 4, 96, 8, nvband=1, seed=9, vcoul_linear=True)
        assert np.array_equal(b["vcoul"], np.arange(96) / 96.0)
        for k in ("aqsmtemp", "aqsntemp", "I_eps_array", "wtilde_array",
                  "inv_igp_index", "indinv", "wx_array"):
            assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("sem", ["gppker", "bgw"])
@pytest.mark.parametrize("shape", [(64, 2, 2048, 8), (9, 3, 100, 7)])
def test_oracle_linear_vcoul_closed_form(sem, shape):
    """Config 1 with vcoul proportional to i: the closed form needs only
    sum_my_igp vcoul[igp] (achtemp, asxtemp) and the gppker achstemp is
    NB |c|^2 c sum vcoul / 2 (gppKer.cpp:96-97, the igp term only)."""
    nb, nv, nc, npg = shape
    ng = nc // npg
    inp = gppk.fill_inputs(1, nb, nc, npg, nvband=nv, vcoul_linear=True)
    vsum = linear_vcoul_sum(nc, ng)
    assert vsum != ng  # not the file's vcoul = 1
    got = orc.sigma(inp, sem)
    ach, asx = closed_form_rows([(3.0, 4.0, 5.0)] * nb, nv, ng, nc, vsum, sem)
    assert rel_err(got["achtemp"], ach) < 1e-12
    assert rel_err(got["asxtemp"], asx) < 1e-12
    if sem == "gppker":
        want = nb * 0.5 * abs(C) ** 2 * C * vsum
        assert abs(got["achstemp"] - want) <= 1e-12 * abs(want)


def masked_problem(k_list):
    """Constant fills, 3 non-occupied bands, wx = {3,4,5} shared, and
    wtilde replaced at the (my_igp, ig) of k_list by w' (value given per
    entry).  Returns the inputs and the per-term closed form: each band
    adds  0.25 vcoul [ sum_all c delw_c(wx) + sum_placed (kept(w') c
    delw_w'(wx) - c delw_c(wx)) ]  with kept() the bgw mask of
    gppkernel.deslippe.09.30.2014.f90:364-366 (always true for gppker)."""
    nb, nc, npg = 3, 64, 8
    inp = gppk.fill_inputs(1, nb, nc, npg, nvband=0)
    w = inp["wtilde_array"].copy()
    for (g, ig, val) in k_list:
        w[g, ig] = val
    inp["wtilde_array"] = w
    return inp


def masked_expected(inp, k_list, sem):
    nb = inp["aqsmtemp"].shape[0]
    ng, nc = inp["I_eps_array"].shape
    wx_rows = np.broadcast_to(inp["wx_array"], (nb, 3)) if inp["wx_array"].ndim == 1 \
        else inp["wx_array"]
    ach = np.zeros(3, complex)
    dropped = 0
    for n1 in range(nb):
        for iw in range(3):
            wx = clamp(wx_rows[n1][iw])
            delw_c = C * np.conj(wx - C) / abs(wx - C) ** 2
            tot = ng * nc * delw_c
            for (_, _, wp) in k_list:
                d = abs(wx - wp) ** 2
                delw = wp * np.conj(wx - wp) / d
                kept = sem != "bgw" or (d > 0.25 and abs(delw) ** 2 < 2.5e5)
                dropped += not kept
                tot += (delw if kept else 0.0) - delw_c
            ach[iw] += 0.25 * C * tot
    return ach, dropped


@pytest.mark.parametrize("sem", ["gppker", "bgw"])
def test_oracle_bgw_mask_drops_exactly_the_window_terms(sem):
    """k = 3 wtilde placed inside the 0.5 window of wx = 4 (|wx - w'|^2 <=
    0.25) and one far away: bgw drops exactly those 3 terms per band (one
    frequency each), gppker keeps them; a mask dropping one term more or
    fewer moves achtemp by ~1e-3 relative."""
    placed = [(0, 5, 4.0 + 0.1j), (3, 17, 4.45 + 0.0j), (7, 63, 3.6 - 0.2j),
              (2, 30, 8.0 + 0.05j)]
    inp = masked_problem(placed)
    got = orc.sigma(inp, sem)["achtemp"]
    want, dropped = masked_expected(inp, placed, sem)
    assert dropped == (3 * 3 if sem == "bgw" else 0)
    assert rel_err(got, want) < 1e-12, (got, want)
    # one dropped term more or fewer is far outside that tolerance
    wp = placed[0][2]
    one = 0.25 * C * (wp * np.conj(4.0 - wp) / abs(4.0 - wp) ** 2)
    assert abs(one) / abs(want[1]) > 1e-4


def test_oracle_bgw_mask_second_condition():
    """A term with |wdiff|^2 > 0.25 but |delw|^2 = |wtilde|^2/|wdiff|^2 >=
    2.5e5 (wtilde = 300, band-resolved wx = 300.55: |wdiff|^2 = 0.3025,
    |delw|^2 = 2.98e5) is dropped by the limitone condition alone."""
    placed = [(1, 9, 300.0 + 0.0j)]
    inp = masked_problem(placed)
    inp["wx_array"] = np.array([[300.55, 4.0, 5.0]] * 3)
    got = orc.sigma(inp, "bgw")["achtemp"]
    want, dropped = masked_expected(inp, placed, "bgw")
    assert dropped == 3  # one per band, at iw = 0
    assert rel_err(got, want) < 1e-12, (got, want)
    keep, _ = masked_expected(inp, placed, "gppker")
    assert rel_err(keep, want) > 1e-3


# The occupied solver's second branch next to the bgw cutoff.
BRANCH2_EPS = (0.01, 0.02, 0.03, 0.05)
BRANCH2_WX = np.array([[-0.2, 3.0, -5.0]])


def branch2_problem():
    """One occupied band, constant fills c = 0.5+0.5i except wtilde =
    0.2 + eps i at four entries, band-resolved wx = (-0.2, 3, -5).  At
    wx = -0.2 those four take the solver's second branch (|wdiff|^2 =
    0.16 + eps^2 <= 0.25) with delw close to -1/2, where cden = 4 wtilde2
    (delw + 1/2) is small and |ssx| = |I_eps delw / (4 delw + 2)| runs from
    10 |I_eps| (eps = 0.01, cut by bgw: wx < 0) to 2 |I_eps| (0.05, kept);
    every other term takes the first branch."""
    inp = gppk.fill_inputs(1, 1, 64, 8, nvband=1)
    w = inp["wtilde_array"].copy()
    for k, eps in enumerate(BRANCH2_EPS):
        w[k, 3 + 7 * k] = 0.2 + eps * 1j
    return dict(inp, wtilde_array=w, wx_array=BRANCH2_WX.copy())


def branch2_expected(sem):
    """asxtemp from the solver written out per term, the second branch as
    -I_eps delw / (4 delw + 2) (the reference's -Omega2 conj(cden) delw /
    |cden|^2 with cden = 4 wtilde2 (delw + 1/2), wtilde2 cancelled)."""
    inp = branch2_problem()
    ng, nc = inp["I_eps_array"].shape
    w = inp["wtilde_array"]
    out = np.zeros(3, complex)
    for iw, wx in enumerate(BRANCH2_WX[0]):
        wd = wx - w
        delw = w / wd
        b1 = (np.abs(wd) ** 2 > 0.25) & (np.abs(delw) ** 2 < 2.5e5)
        if sem == "bgw":
            ssx1 = w ** 2 * C * np.conj(wx * wx - w ** 2) / np.abs(wx * wx - w ** 2) ** 2
        else:
            ssx1 = w ** 2 * C / wx ** 2
        ssx2 = -C * delw / (4 * delw + 2)
        ssx = np.where(b1, ssx1, np.where(np.abs(delw) ** 2 > 1e-6, ssx2, 0))
        if sem == "bgw" and wx < 0:
            ssx = np.where(np.abs(ssx) > 4 * abs(C), 0, ssx)
        out[iw] = np.sum(ssx) * abs(C) ** 2 * (-1 if sem == "bgw" else 1)
    return out


@pytest.mark.parametrize("sem", ["gppker", "bgw"])
def test_oracle_second_branch_next_to_cutoff(sem):
    inp = branch2_problem()
    got = orc.sigma(inp, sem)["asxtemp"]
    want = branch2_expected(sem)
    assert rel_err(got, want) <= 1e-12, (got, want)
    # the four entries really are in the second branch, two of them past
    # the cutoff
    w = inp["wtilde_array"]
    wd = -0.2 - w
    delw = w / wd
    sel = np.abs(wd) ** 2 <= 0.25
    assert sel.sum() == 4
    r = np.abs(delw / (4 * delw + 2))[sel]
    assert (r > 4).sum() == 2 and (r < 4).sum() == 2
