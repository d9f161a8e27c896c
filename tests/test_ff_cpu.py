"""CPU tests of the full-frequency (FF) path: the oracle against the
reference's own FF code (the unmodified mini-app binary and its kernel
functions on random inputs), closed forms of the mini-app's fills, the
frequency-grid search, invariants, the input generator and the CLI."""
import json
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, gppk, gpp_oracle as orc


def c(a):
    a = np.asarray(a, np.float64)
    return a[..., 0] + 1j * a[..., 1]


def rel(a, b):
    """max |a - b| / max |b| over the whole array (0 if both vanish)."""
    a = np.atleast_1d(np.asarray(a, np.complex128))
    b = np.atleast_1d(np.asarray(b, np.complex128))
    s = np.max(np.abs(b))
    d = np.max(np.abs(a - b))
    return d if s == 0 else d / s


def kat_problem(r):
    return dict(number_bands=r["nb"], nvband=r["nv"], ncouls=r["nc"],
                ngpown=r["ng"], nFreq=r["nf"], nfreqeval=r["ne"],
                aqsmtemp=c(r["aqsm"]), aqsntemp=c(r["aqsn"]),
                I_epsR_array=c(r["er"]), I_epsA_array=c(r["ea"]),
                vcoul=np.array(r["vcoul"]),
                inv_igp_index=np.array(r["inv"], np.int32),
                indinv=np.array(r["indinv"], np.int32), ekq=np.array(r["ekq"]),
                dFreqGrid=np.array(r["grid"]), pref=np.array(r["pref"]),
                dFreqBrd=c(r["brd"]), schDt_matrix=c(r["schdt"]),
                freqevalmin=r["fmin"], freqevalstep=r["fstep"],
                pref_zb=r["pref_zb"], occ=r["occ"])


KATS = json.loads((GOLDEN / "ref_ff_kat.json").read_text())
CONST = json.loads((GOLDEN / "ref_ff_constfill.json").read_text())


@pytest.mark.parametrize("r", KATS, ids=lambda r: f"seed{r['seed']}")
def test_oracle_matches_reference_kernels(r):
    """ffKernel.cpp's own achsDtemp_Kernel/asxDtemp_Kernel/achDtemp_Kernel/
    achDtemp_cor_Kernel on random inputs (non-uniform grid, wx below zero,
    inside and beyond the grid, strided and permuted index maps)."""
    got = orc.ff_sigma(kat_problem(r), "cpp")
    assert abs(got["achsDtemp"] - complex(*r["achsDtemp"])) <= 1e-13 * abs(complex(*r["achsDtemp"]))
    for k in ("asxDtemp", "achDtemp", "achDtemp_cor"):
        assert rel(got[k], c(r[k])) <= 1e-13, k


def test_kat_inputs_exercise_both_cor_branches():
    """At least one fixture has a non-zero achDtemp_cor, and the band energies
    put wx on both sides of zero."""
    assert any(np.any(c(r["achDtemp_cor"]) != 0) for r in KATS)
    for r in KATS:
        wx = r["fmin"] - np.array(r["ekq"]) + r["fstep"]
        assert (wx > 0).any() and (wx <= 0).any()


def closed_form(sem, nb, nv, nc, ng, nf, ne, beta=0.0, gamma=0.0):
    """The mini-app's fills (aqsm = aqsn = 0.5+0.5i, vcoul = 1, identity
    maps, ekq = -10+n1, schDt = 0, grid g_f = 2f) with I_eps linear in the
    grid frequency and constant over (G', G):
        I_epsR[f] = 0.5+0.5i + beta g_f,   I_epsA[f] = 0.5-0.5i + gamma g_f
    (beta = gamma = 0: the fills themselves).  Every interpolation of both
    sources evaluates the linear interpolant at x = |wx| with weights that
    reproduce a linear function exactly, inside the grid or extrapolated
    from the interval a search falls back to; the cor/ach2 weights carry a
    factor -1/2.  With aqsn*conj(aqsm) = 1/2 each sum is K/2 times the
    interpolated value, K = ng*nc."""
    K = ng * nc
    IR = lambda x: 0.5 + 0.5j + beta * x
    IA = lambda x: 0.5 - 0.5j + gamma * x
    g_achs = 2.0 * (1 if sem == "cpp" else 0)  # cpp: slice 1, f90: slice 0
    achs = -0.5 * nb * 0.5 * K * IR(g_achs)
    asx = np.zeros(ne, complex)
    cor = np.zeros(ne, complex)
    a2 = np.zeros(ne, complex)
    for n1 in range(nb):
        occ = n1 < nv
        for iw in range(ne):
            wx = 0.0 - (-10.0 + n1) + (0.5 if sem == "cpp" else iw * 0.5)
            x = abs(wx)
            if occ:
                src = IR if (sem == "cpp" or wx >= 0) else IA
                asx[iw] += 0.5 * K * src(x)
            pos = wx > 0 if sem == "cpp" else wx >= 0
            s = 0.5 * K * -0.5 * (IR(x) - IA(x))
            if pos:
                if sem == "f90" or not occ:
                    a2[iw] += s
                if not occ:
                    cor[iw] += s
            elif occ:
                cor[iw] += s
    return achs, asx, cor, a2


def linear_eps(inp, beta, gamma):
    """Config-1 inputs with I_eps linear in the grid frequency."""
    g = inp["dFreqGrid"][:, None, None]
    er = np.broadcast_to(0.5 + 0.5j + beta * g, inp["I_epsR_array"].shape)
    ea = np.broadcast_to(0.5 - 0.5j + gamma * g, inp["I_epsA_array"].shape)
    return dict(inp, I_epsR_array=np.ascontiguousarray(er),
                I_epsA_array=np.ascontiguousarray(ea))


LINEAR_EPS = (0.3 - 0.2j, -0.15 + 0.05j)


SHAPES = [(4, 2, 64, 8, 6, 5), (14, 13, 40, 5, 9, 3), (3, 0, 30, 3, 4, 2),
          (12, 3, 50, 6, 4, 4), (16, 14, 24, 4, 12, 30)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("sem", ["cpp", "f90"])
def test_oracle_closed_form(shape, sem):
    nb, nv, nc, ng, nf, ne = shape
    inp = gppk.ff_fill_inputs(1, nb, nv, nc, ng, nf, ne, semantics=sem)
    got = orc.ff_sigma(inp, sem)
    achs, asx, cor, a2 = closed_form(sem, *shape)
    assert abs(got["achsDtemp"] - achs) <= 1e-12 * abs(achs)
    assert rel(got["asxDtemp"], asx) <= 1e-12
    assert rel(got["achDtemp_cor"], cor) <= 1e-12
    assert rel(got["ach2Dtemp"], a2) <= 1e-12
    assert np.all(got["achDtemp"] == 0) and np.all(got["achDtemp_corb"] == 0)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("sem", ["cpp", "f90"])
def test_oracle_linear_in_frequency_closed_form(shape, sem):
    """I_eps linear in the grid frequency: pins which two slices each
    interpolation reads and with which weights (a swapped pair, a weight on
    the wrong slice or the wrong sign of wx changes the value), the slice of
    achsDtemp, and I_epsA below zero frequency."""
    inp = linear_eps(gppk.ff_fill_inputs(1, *shape, semantics=sem), *LINEAR_EPS)
    got = orc.ff_sigma(inp, sem)
    achs, asx, cor, a2 = closed_form(sem, *shape, *LINEAR_EPS)
    assert abs(got["achsDtemp"] - achs) <= 1e-12 * abs(achs)
    assert rel(got["asxDtemp"], asx) <= 1e-12
    assert rel(got["achDtemp_cor"], cor) <= 1e-12
    assert rel(got["ach2Dtemp"], a2) <= 1e-12


@pytest.mark.parametrize("r", CONST, ids=lambda r: "{number_bands}-{nvband}-{ncouls}-{ngpown}-{nFreq}-{nfreqeval}".format(**r))
def test_oracle_matches_reference_binary_constfill(r):
    """The lines the unmodified ffKernel.cpp prints (its main hands asxDtemp
    to achDtemp_Kernel, so the printed asxDtemp includes achDtemp)."""
    shape = [r[k] for k in ("number_bands", "nvband", "ncouls", "ngpown", "nFreq", "nfreqeval")]
    inp = gppk.ff_fill_inputs(1, *shape, semantics="cpp")
    got = orc.ff_sigma(inp, "cpp")
    assert abs(got["achsDtemp"] - complex(*r["achsDtemp"])) <= 1e-6
    assert abs(got["asxDtemp"][0] + got["achDtemp"][0] - complex(*r["asxDtemp"])) <= 1e-6
    assert abs(got["achDtemp_cor"][0] - complex(*r["achDtemp_cor"])) <= 1e-6
    achs, asx, cor, _ = closed_form("cpp", *shape)
    assert abs(achs - complex(*r["achsDtemp"])) <= 1e-6
    assert abs(asx[0] - complex(*r["asxDtemp"])) <= 1e-6
    assert abs(cor[0] - complex(*r["achDtemp_cor"])) <= 1e-6


# ---- ffKernel.f90 frequency-grid terms (:335-411) on inputs whose sums
# have closed forms.  aqsm = 0 removes every I_eps term, so the outputs are
#   achDtemp      = sum_n1 sum_{f < nF-1} -i pref[f] S[n1][f] / (wx - g[f] + brd[f])
#   achDtemp_corb = the same with the pole at -g[f] - brd[f] for occupied n1
#   ach2Dtemp     = sum_n1 sum_{f >= 1} -i pref_zb avg_f it_f,
#                   it_f = -log|(wx - g[f]) / (wx - g[f-1])|
#   achDtemp_cor  = sum_n1 sum_{f >= 1} -i pref_zb ((S[f-1] + lin_f/dg_f a) it'_f + lin_f)
# and the grid, band energies and evaluation frequencies keep every ratio
# away from the 1e-4 / 1e4 clamps, so log sums telescope.
GRID = np.array([0.0, 1.0, 2.5, 4.0, 6.0])
GRID_BRD = np.array([0.3 + 0.2j, -0.1 + 0.4j, 0.2 + 0.1j, 0.05 + 0.3j, 0.1 + 0.25j])
GRID_PREF = np.array([0.7, 1.1, 0.9, 1.3, 0.6])
GRID_EKQ = np.array([-12.0, -11.0, -3.2])   # band 0 occupied; band 2's wx
                                             # runs through the grid
GRID_C = np.array([0.8 - 0.3j, -0.5 + 1.1j, 0.25 + 0.6j])


def grid_problem(schdt):
    nb, nv, nc, ng, nf, ne = 3, 1, 16, 2, 5, 6
    inp = gppk.ff_fill_inputs(2, nb, nv, nc, ng, nf, ne, semantics="f90", seed=11)
    return dict(inp, aqsmtemp=np.zeros_like(inp["aqsmtemp"]), dFreqGrid=GRID.copy(),
                dFreqBrd=GRID_BRD.copy(), pref=GRID_PREF.copy(), ekq=GRID_EKQ.copy(),
                schDt_matrix=np.ascontiguousarray(schdt, np.complex128),
                freqevalmin=0.3, freqevalstep=0.7, pref_zb=0.45)


def grid_wx(inp):
    iw = np.arange(inp["nfreqeval"])
    return [inp["freqevalmin"] - e + iw * inp["freqevalstep"] for e in inp["ekq"]]


GRID_CASE_NAMES = ("delta0", "delta2", "delta4", "constant", "linear")


def grid_case(name):
    """(name, inputs, {output: expected}); built when a test runs, so that
    collecting the tests loads no library."""
    return {c[0]: c for c in grid_cases()}[name]


def grid_cases():
    """(name, inputs, {output: expected}) for single-point, constant and
    linear S(f)."""
    g, L, pz = GRID, len(GRID) - 1, 0.45
    cases = []
    lg = lambda x: np.log(np.abs(x))
    for k in (0, 2, L):  # S = c delta_{f,k}
        S = np.zeros((3, len(g)), complex)
        S[:, k] = GRID_C
        inp = grid_problem(S)
        wxs = grid_wx(inp)
        ach = sum(-1j * GRID_PREF[k] * c / (wx - g[k] + GRID_BRD[k]) for c, wx in zip(GRID_C, wxs))
        corb = sum(-1j * GRID_PREF[k] * c / (wx - (-g[k] if n1 == 0 else g[k]) + GRID_BRD[k])
                   for n1, (c, wx) in enumerate(zip(GRID_C, wxs)))
        if k == L:  # the last grid point has no pole term
            ach = corb = np.zeros(len(wxs[0]), complex)
        lo, hi = max(k - 1, 0), min(k + 1, L)
        a2 = sum(1j * pz * (c / 2) * (lg(wx - g[hi]) - lg(wx - g[lo])) for c, wx in zip(GRID_C, wxs))
        cases.append((f"delta{k}", inp, {"achDtemp": ach, "achDtemp_corb": corb, "ach2Dtemp": a2}))
    # S constant per band: lin = 0; the log sums telescope
    S = np.repeat(GRID_C[:, None], len(g), axis=1)
    inp = grid_problem(S)
    wxs = grid_wx(inp)
    a2 = sum(1j * pz * c * (lg(wx - g[L]) - lg(wx - g[0])) for c, wx in zip(GRID_C, wxs))
    cor = sum((-1j * pz * c * (lg(wx + g[L]) - lg(wx + g[0])) if n1 == 0 else
               1j * pz * c * (lg(wx - g[L]) - lg(wx - g[0])))
              for n1, (c, wx) in enumerate(zip(GRID_C, wxs)))
    cases.append(("constant", inp, {"ach2Dtemp": a2, "achDtemp_cor": cor}))
    # S = b g (linear in the grid frequency): lin3_f = b w it'_f + b dg_f with
    # w = wx (unoccupied; ffKernel.f90:404) or -freqevalmin - ekq + iw step
    # (occupied; :402), so the sum is b w (sum it') + b (g_L - g_0)
    S = GRID_C[:, None] * g[None, :]
    inp = grid_problem(S)
    wxs = grid_wx(inp)
    iw = np.arange(inp["nfreqeval"])
    cor = 0
    for n1, (b, wx) in enumerate(zip(GRID_C, wxs)):
        if n1 == 0:
            w = -inp["freqevalmin"] - inp["ekq"][0] + iw * inp["freqevalstep"]
            s_it = lg(wx + g[L]) - lg(wx + g[0])
        else:
            w = wx
            s_it = -(lg(wx - g[L]) - lg(wx - g[0]))
        cor = cor - 1j * pz * (b * w * s_it + b * (g[L] - g[0]))
    cases.append(("linear", inp, {"achDtemp_cor": cor}))
    return cases


@pytest.mark.parametrize("name", GRID_CASE_NAMES)
def test_oracle_f90_grid_terms_closed_forms(name):
    """The oracle's ffKernel.f90 frequency-grid terms against their closed
    forms: single-point S pins each pole term, wx (iw, not iw-1), the
    occupied cor pole, the exclusion of the last grid point from the pole
    sums and the half-sum avg; constant and linear S pin the log ratios, the
    left-point expansion of lin3 and its + lin term."""
    name, inp, want = grid_case(name)
    got = orc.ff_sigma(inp, "f90")
    assert got["achsDtemp"] == 0 and np.all(got["asxDtemp"] == 0)
    for k, v in want.items():
        assert rel(got[k], v) <= 1e-13, (name, k, got[k], v)


def test_frequency_search_quirks():
    g = 2.0 * np.arange(6)  # 0 2 4 6 8 10
    # C++ loop 1: strict search; a hit in the first interval, a grid point
    # and a point beyond the grid all fall back to the last interval.
    assert orc.ff_fact("cpp", 1, False, 5.0, g) == (2, 0.5, 0.5)
    assert orc.ff_fact("cpp", 1, False, 1.0, g)[0] == 4
    assert orc.ff_fact("cpp", 1, False, 4.0, g)[0] == 4
    f, f1, f2 = orc.ff_fact("cpp", 1, False, 10.5, g)
    assert f == 4 and f1 + f2 == pytest.approx(1.0) and f1 < 0
    f, f1, f2 = orc.ff_fact("cpp", 1, False, -5.0, g)  # uses -wx
    assert (f, f1, f2) == (2, 0.5, 0.5)
    # C++ loop 2, wx > 0: no fallback, weights sum to -1/2
    f, f1, f2 = orc.ff_fact("cpp", 2, False, 1.0, g)
    assert f == 0 and f1 + f2 == pytest.approx(-0.5)
    assert orc.ff_fact("cpp", 2, True, 3.0, g)[0] == -1   # occupied, wx > 0
    assert orc.ff_fact("cpp", 2, False, -3.0, g)[0] == -1  # unoccupied, wx <= 0
    # F90: half-open search, last interval when nothing matches
    assert orc.ff_fact("f90", 1, True, 4.0, g)[0] == 2
    assert orc.ff_fact("f90", 1, True, 0.0, g)[0] == 0
    assert orc.ff_fact("f90", 1, True, 11.0, g)[0] == 4
    assert orc.ff_fact("f90", 1, False, 4.0, g)[0] == -1  # asx: occupied only
    assert orc.ff_fact("f90", 2, False, -1.0, g)[0] == -1


def test_linearity_in_schdt_and_aqsn():
    inp = gppk.ff_fill_inputs(2, 5, 2, 40, 5, 6, 7, semantics="f90", seed=9)
    base = orc.ff_sigma(inp, "f90")
    two = orc.ff_sigma(dict(inp, schDt_matrix=2 * inp["schDt_matrix"]), "f90")
    for k in ("achDtemp", "achDtemp_corb"):
        assert rel(two[k], 2 * base[k]) <= 1e-13, k
    three = orc.ff_sigma(dict(inp, aqsntemp=3 * inp["aqsntemp"]), "f90")
    for k in ("asxDtemp",):
        assert rel(three[k], 3 * base[k]) <= 1e-13, k
    assert abs(three["achsDtemp"] - 3 * base["achsDtemp"]) <= 1e-13 * abs(base["achsDtemp"])


def test_band_additivity_of_grid_terms():
    """achDtemp is a sum over bands: splitting the bands splits it."""
    inp = gppk.ff_fill_inputs(2, 6, 3, 30, 3, 5, 4, semantics="cpp", seed=4)
    full = orc.ff_sigma(inp, "cpp")["achDtemp"]
    parts = []
    for a, b in ((0, 2), (2, 6)):
        d = dict(inp, number_bands=b - a, nvband=max(0, min(3, b) - a),
                 aqsmtemp=inp["aqsmtemp"][a:b], aqsntemp=inp["aqsntemp"][a:b],
                 ekq=inp["ekq"][a:b], schDt_matrix=inp["schDt_matrix"][a:b])
        parts.append(orc.ff_sigma(d, "cpp")["achDtemp"])
    assert rel(parts[0] + parts[1], full) <= 1e-13


def test_generator_is_deterministic_and_shaped():
    a = gppk.ff_fill_inputs(2, 4, 2, 48, 6, 5, 3, seed=1)
    b = gppk.ff_fill_inputs(2, 4, 2, 48, 6, 5, 3, seed=1)
    for k in gppk.FF_ARRAYS:
        assert np.array_equal(a[k], b[k]), k
    assert a["I_epsR_array"].shape == (5, 6, 48)
    assert np.all(np.diff(a["dFreqGrid"]) > 0) and np.all(np.diff(a["ekq"]) >= 0)
    assert list(a["inv_igp_index"]) == [0, 8, 16, 24, 32, 40]
    c1 = gppk.ff_fill_inputs(1, 4, 2, 48, 6, 5, 3, semantics="f90")
    assert np.all(c1["dFreqBrd"] == 0.1j) and np.all(c1["I_epsA_array"] == 0.5 - 0.5j)


def test_cli_usage():
    out = subprocess.run([str(gppk.FF_CLI_PATH)], capture_output=True, text=True)
    assert out.returncode == 0
    assert "number_bands nvband ncouls ngpown nFreq nfreqeval" in out.stdout


LOG_CHECK = r"""
#include "ff_math.h"
#include <cmath>
#include <cstdio>
#include <random>
using namespace gpp::ff;
// |got - want| in units of ulp(max(|want|, ln 2)): the absolute error that a
// difference of two logarithms inherits
static double units(double got, long double want) {
  const double w = std::fmax(std::fabs((double)want), 0.6931471805599453);
  const double u = std::nextafter(w, INFINITY) - w;
  return (double)(std::fabs((long double)got - want) / u);
}
static long double clampl(long double r) {
  return r < 1e-4L ? (long double)1e-4 : r > 1e4L ? (long double)1e4 : r;
}
int main() {
  std::mt19937_64 g(7);
  std::uniform_real_distribution<double> wide(std::log(1e-300), std::log(1e300));
  std::uniform_real_distribution<double> near(-1.0, 1.0);
  std::uniform_real_distribution<double> u01(0.0, 1.0);
  double w = 0, n = 0, sub = 0, d = 0;  // d: INFINITY once a difference fails
  for (int i = 0; i < 2000000; ++i) {
    const double x = std::exp(wide(g)) * (i & 1 ? -1 : 1);
    w = std::fmax(w, units(log_abs(x, kLogTable), logl(fabsl((long double)x))));
    const double y = 1.0 + near(g) * std::pow(10.0, -1 - i % 15);
    n = std::fmax(n, units(log_abs(y, kLogTable), logl((long double)y)));
    const double s = std::ldexp(u01(g), -1022 - (i % 52));
    if (s != 0) sub = std::fmax(sub, units(log_abs(s, kLogTable), logl((long double)s)));
    // the interval term: clamp(log|a| - log|b|) against log(clamp|a/b|),
    // a and b as |wx - zb| at adjacent grid points (ratios near 1 to far)
    const double b = std::exp(wide(g) * 0.02);
    const double a = b * (1.0 + near(g) * std::pow(10.0, 4 - i % 12));
    const double t = clamp_log_ratio(log_abs(a, kLogTable) - log_abs(b, kLogTable));
    const long double want = logl(clampl(fabsl((long double)a / b)));
    const double big = std::fmax(std::fabs(std::log(std::fabs(a))), std::fabs(std::log(b)));
    // two logarithms (1.3 units each), the subtraction's rounding and the
    // bounds' own rounding: 3 ulp of the larger logarithm plus 1 of t
    const double ub = std::fmax(big, 0.7), ut = std::fabs(t);
    if (!(std::fabs((long double)t - want) <=
          3.0 * (std::nextafter(ub, INFINITY) - ub) + (std::nextafter(ut, INFINITY) - ut)))
      d = INFINITY;
  }
  // the kernel's recurrence (IntervalLog) against the f90's own steps,
  // intfact = clamp(|a/b|) and log(intfact) in double, over grids with
  // wx on grid points, repeated points, huge and non-finite ones
  int seq_ok = 1;
  const double grid[] = {0.0, 0.5, 1.0, 1.0, 1.5, 2.0, 2.0 + 1e-9, 3.0, 3.0 + 1e-14,
                         7.25, 1e300, INFINITY, 4.0, -2.0, NAN, 5.0, 6.0};
  const int ng = sizeof grid / sizeof grid[0];
  const double wxs[] = {1.0, 2.0, 0.25, -3.0, 3.0, 7.25, 2.0 + 5e-10};
  for (double wx : wxs)
    for (int sgn = -1; sgn <= 1; sgn += 2) {
      IntervalLog st;
      st.start(wx + sgn * grid[0], kLogTable);
      for (int f = 1; f < ng; ++f) {
        const double a = wx + sgn * grid[f], b = wx + sgn * grid[f - 1];
        const double t = st.step(a, kLogTable);
        double ratio = std::fabs(a / b);
        if (ratio < 1e-4) ratio = 1e-4;
        if (ratio > 1e4) ratio = 1e4;
        const double ref = std::log(ratio);
        if (std::isnan(ref)) {
          seq_ok &= std::isnan(t);
          continue;
        }
        const double big = std::fmax(std::fmax(std::fabs(std::log(std::fabs(a))),
                                               std::fabs(std::log(std::fabs(b)))), 0.7);
        const double tol = std::isfinite(big)
            ? 3.0 * (std::nextafter(big, INFINITY) - big) + 0x1p-50 : 0.0;
        seq_ok &= std::fabs(t - ref) <= tol;
      }
    }
  const double lo = clamp_log_ratio(log_abs(0.0, kLogTable) - log_abs(2.0, kLogTable));
  const double hi = clamp_log_ratio(log_abs(3.0, kLogTable) - log_abs(-0.0, kLogTable));
  const double nan00 = clamp_log_ratio(log_abs(0.0, kLogTable) - log_abs(0.0, kLogTable));
  std::printf("%.6f %.6f %.6f %d\n", w, n, sub, (int)(d < INFINITY));
  std::printf("%d %d %d %d %d %d %d %d\n", seq_ok,
              log_abs(0.0, kLogTable) == -INFINITY && log_abs(-0.0, kLogTable) == -INFINITY,
              log_abs(INFINITY, kLogTable) == INFINITY && log_abs(-INFINITY, kLogTable) == INFINITY,
              (int)std::isnan(log_abs(NAN, kLogTable)),
              std::fabs(log_abs(1.0, kLogTable)) <= 0x1p-54 && log_abs(-1.0, kLogTable) == log_abs(1.0, kLogTable),
              lo == std::log(1e-4) && hi == std::log(1e4),
              (int)std::isnan(nan00),
              kLogRatioLo == std::log(1e-4) && kLogRatioHi == std::log(1e4));
}
"""


def test_grid_kernel_logarithm_accuracy(tmp_path):
    """The logarithm of the frequency-grid kernel (csrc/ff_math.h, built for
    the host with g++) against long double logl, as an absolute error in
    units of ulp(max(|log x|, ln 2)): at most 1.3 over |x| in [1e-300,
    1e300] and for subnormals, at most 0.6 within 0.1 of 1; the clamped
    difference of two logarithms (the interval term) within 3 ulp of the
    larger logarithm plus 1 ulp of the result of log(clamp(|a/b|)); special
    values (log 0 = -inf, log inf = inf, NaN, |log 1| <= 2^-54, 0/x and x/0
    clamped to the bounds, 0/0 NaN as the reference's ratio) and the clamp
    bounds equal to log(1e-4) and log(1e4) as double; and the kernel's
    recurrence over consecutive grid points (IntervalLog: one logarithm per
    point, exact form only for differences outside the bounds or a zero
    end) against the f90's clamp(|a/b|) and log over grids holding zeros,
    repeated points and non-finite values."""
    import shutil
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    src = tmp_path / "log_check.cpp"
    src.write_text(LOG_CHECK)
    exe = tmp_path / "log_check"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-I",
                    str(gppk.PKG_DIR / "csrc"), str(src), "-o", str(exe)], check=True)
    line1, line2 = subprocess.run([str(exe)], capture_output=True, text=True,
                                  check=True, timeout=120).stdout.splitlines()
    w, n, sub, diff_ok = line1.split()
    assert float(w) <= 1.3 and float(sub) <= 1.3 and float(n) <= 0.6, line1
    assert diff_ok == "1"
    assert line2.split() == ["1"] * 8, line2


def test_log_table_is_generated():
    """csrc/ff_log_table.h is what csrc/ff_log_table.py writes (table from
    exact rational and 50-digit decimal arithmetic, each entry rounded
    once), and its polynomial's economization bound is below 3.2e-13."""
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "ff_log_table", gppk.PKG_DIR / "csrc" / "ff_log_table.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    assert (gppk.PKG_DIR / "csrc" / "ff_log_table.h").read_text() == mod.header()
    _, dropped = mod.economized_q()
    assert float(dropped) < 3.2e-13
    # every entry: invc the double nearest 1/c, logc the double nearest
    # -ln(invc) (checked here with math.log on the exact invc, to 1 ulp)
    import math
    for i, (invc, logc) in enumerate(mod.table()):
        c = 0.5 + (2 * i + 1) / 1024
        assert invc == 1 / c
        assert abs(logc + math.log(invc)) <= math.ulp(abs(logc))
