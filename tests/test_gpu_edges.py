"""GPU parity on the edges the oracle's closed forms pin: the frequency
clamp of gppKer.cpp:279 (|wx| < 1e-6 -> 1e-6), vcoul proportional to i, the
bgw mask on a known subset of terms (gppkernel.deslippe.09.30.2014.f90:
364-366), and concurrent drop-in calls from two host threads."""
import ctypes
import threading

import numpy as np
import pytest

from conftest import gppk, gpp_oracle as orc, rel_err
from test_oracle_cpu import (CLAMP_ROWS, closed_form_rows, linear_vcoul_sum,
                             masked_expected, masked_problem)

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-11


def say(*a):
    print(*a, flush=True)


def check(got, ref, tol=FP64_TOL, keys=("achtemp", "asxtemp")):
    for k in keys:
        e = rel_err(got[k], ref[k])
        assert e <= tol, (k, e, got[k], ref[k])


@pytest.mark.parametrize("sem", ["gppker", "strict", "bgw"])
@pytest.mark.parametrize("nv", [0, 2])
def test_clamp_band_resolved(handle, sem, nv):
    nb, nc, npg = len(CLAMP_ROWS), 64, 8
    say("clamp band-resolved", sem, nv)
    inp = gppk.fill_inputs(1, nb, nc, npg, nvband=nv)
    inp["wx_array"] = np.array(CLAMP_ROWS, np.float64)
    ref = orc.sigma(inp, sem)
    ach, asx = closed_form_rows(CLAMP_ROWS, nv, nc // npg, nc, nc // npg,
                                "bgw" if sem == "bgw" else "gppker")
    got = handle.sigma(inp, semantics=sem)
    check(got, ref)
    assert rel_err(got["achtemp"], ach) < 1e-12
    if nv:
        assert rel_err(got["asxtemp"], asx) < 1e-12
    # the batched launch takes the same clamp
    b = handle.batch([inp, inp], semantics=sem)
    for r in b:
        check(r, ref)
    if sem != "bgw":
        m = handle.sigma(inp, precision="mixed", semantics=sem)
        check(m, ref, tol=1e-6)


@pytest.mark.parametrize("sem", ["gppker", "bgw"])
@pytest.mark.parametrize("wx", [(0.0, -5e-7, 4.0), (5e-7, -1e-6, 1e-6)])
def test_clamp_shared(handle, sem, wx):
    say("clamp shared", sem, wx)
    inp = gppk.fill_inputs(1, 5, 256, 8, nvband=2)
    inp["wx_array"] = np.array(wx)
    ref = orc.sigma(inp, sem)
    for prec in ("fp64", "fp64_factored"):
        check(handle.sigma(inp, precision=prec, semantics=sem), ref)
    # random inputs through the drop-in, which takes the band sum first
    inp2 = gppk.fill_inputs(2, 6, 512, 8, nvband=2, seed=4)
    inp2["wx_array"] = np.array(wx)
    ref2 = orc.sigma(inp2, sem)
    for prec in ("fp64", "fp64_factored"):
        check(handle.sigma(inp2, precision=prec, semantics=sem), ref2)


@pytest.mark.parametrize("sem", ["gppker", "bgw"])
def test_linear_vcoul(handle, sem):
    nb, nv, nc, npg = 64, 2, 2048, 8
    say("linear vcoul", sem)
    inp = gppk.fill_inputs(1, nb, nc, npg, nvband=nv, vcoul_linear=True)
    dev = gppk.fill_inputs_device(1, nb, nc, npg, nvband=nv, vcoul_linear=True)
    assert np.array_equal(dev["vcoul"].cpu().numpy(), inp["vcoul"])
    vsum = linear_vcoul_sum(nc, nc // npg)
    ach, asx = closed_form_rows([(3.0, 4.0, 5.0)] * nb, nv, nc // npg, nc,
                                vsum, sem)
    ref = orc.sigma(inp, sem)
    for prec in ("fp64", "fp64_factored"):
        got = handle.sigma(dev, precision=prec, semantics=sem)
        check(got, ref)
        assert rel_err(got["achtemp"], ach) < 1e-12
        assert rel_err(got["asxtemp"], asx) < 1e-12
    # random inputs with the linear factor
    r = gppk.fill_inputs(3, 12, 700, 7, nvband=3, seed=2, vcoul_linear=True)
    check(handle.sigma(r, semantics=sem), orc.sigma(r, sem))


def test_cli_linear_vcoul_prints_closed_form():
    import subprocess
    out = subprocess.run([str(gppk.CLI_PATH), "64", "2", "2048", "8",
                          "--vcoul=linear"], capture_output=True, text=True,
                         check=True).stdout
    vsum = linear_vcoul_sum(2048, 256)
    ach, asx = closed_form_rows([(3.0, 4.0, 5.0)] * 64, 2, 256, 2048, vsum)
    for iw in range(3):
        line = next(l for l in out.splitlines() if l.startswith(f"achtemp[{iw}]"))
        re_, im_ = line.split("=")[1].strip().strip("()").split(",")
        got = complex(float(re_), float(im_))
        assert abs(got - ach[iw]) <= 1e-13 * abs(ach[iw]), (line, ach[iw])


def test_cli_batch_and_tune():
    import subprocess
    out = subprocess.run([str(gppk.CLI_PATH), "6", "2", "512", "8", "--config=2",
                          "--batch=3", "--tune", "--precision=fp64"],
                         capture_output=True, text=True, check=True).stdout
    assert "tuned tile:" in out
    for b in range(3):
        inp = gppk.fill_inputs(2, 6, 512, 8, nvband=2, seed=12345 + b)
        ref = orc.sigma(inp)
        for iw in range(3):
            line = next(l for l in out.splitlines()
                        if l.startswith(f"[{b}] achtemp[{iw}]"))
            re_, im_ = line.split("=")[1].strip().strip("()").split(",")
            got = complex(float(re_), float(im_))
            assert abs(got - ref["achtemp"][iw]) <= 1e-13 * abs(ref["achtemp"][iw]) + 1e-300


@pytest.mark.parametrize("sem", ["gppker", "strict", "bgw"])
def test_bgw_mask_subset(handle, sem):
    placed = [(0, 5, 4.0 + 0.1j), (3, 17, 4.45 + 0.0j), (7, 63, 3.6 - 0.2j),
              (2, 30, 8.0 + 0.05j)]
    say("bgw mask subset", sem)
    inp = masked_problem(placed)
    ref = orc.sigma(inp, sem)
    want, _ = masked_expected(inp, placed, "bgw" if sem == "bgw" else "gppker")
    for prec in ("fp64", "fp64_factored"):
        got = handle.sigma(inp, precision=prec, semantics=sem)
        check(got, ref, keys=("achtemp",))
        assert rel_err(got["achtemp"], want) < 1e-12
    for r in handle.batch([inp, inp], semantics=sem):
        check(r, ref, keys=("achtemp",))
    # the limitone condition alone, band-resolved frequencies
    p2 = [(1, 9, 300.0 + 0.0j)]
    inp2 = masked_problem(p2)
    inp2["wx_array"] = np.array([[300.55, 4.0, 5.0]] * 3)
    want2, _ = masked_expected(inp2, p2, "bgw" if sem == "bgw" else "gppker")
    got2 = handle.sigma(inp2, semantics=sem)
    check(got2, orc.sigma(inp2, sem), keys=("achtemp",))
    assert rel_err(got2["achtemp"], want2) < 1e-12


def test_dropin_from_two_threads():
    """gpp_sigma from two host threads at once (different inputs): the
    shared default handle serialises them and each gets its own result."""
    lib = gppk.load_library()
    inputs = [gppk.fill_inputs(2, 6 + i, 640, 8, nvband=2, seed=40 + i)
              for i in range(2)]
    refs = [orc.sigma(x) for x in inputs]
    start = threading.Barrier(2)
    results = [None, None]

    def run(i):
        x = inputs[i]
        ach = (gppk.c128 * 3)()
        asx = (gppk.c128 * 3)()
        P = lambda k: x[k].ctypes.data
        start.wait()
        out = []
        for _ in range(5):
            rc = lib.gpp_sigma(x["number_bands"], 2, x["ncouls"], x["ngpown"],
                               P("aqsmtemp"), P("aqsntemp"), P("I_eps_array"),
                               P("wtilde_array"), P("vcoul"), P("inv_igp_index"),
                               P("indinv"), P("wx_array"), ctypes.byref(ach),
                               ctypes.byref(asx))
            out.append((rc, np.array([complex(z.re, z.im) for z in ach]),
                        np.array([complex(z.re, z.im) for z in asx])))
        results[i] = out

    th = [threading.Thread(target=run, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for i in range(2):
        for rc, ach, asx in results[i]:
            assert rc == 0
            check({"achtemp": ach, "asxtemp": asx}, refs[i])


@pytest.mark.parametrize("sem", ["gppker", "strict", "bgw"])
def test_second_branch_next_to_cutoff(handle, sem):
    """The occupied solver's second branch, evaluated on the GPU as
    -I_eps delw / (4 delw + 2), against the oracle's form of the reference
    (-Omega2 conj(cden) delw / |cden|^2) and the closed form, with |ssx|
    from 2 to 10 |I_eps| around the bgw cutoff (4 |I_eps|, wx < 0)."""
    from test_oracle_cpu import branch2_expected, branch2_problem
    inp = branch2_problem()
    ref = orc.sigma(inp, sem)
    for prec in ("fp64", "mixed"):
        got = handle.sigma(inp, precision=prec, semantics=sem)
        check(got, ref, keys=("asxtemp",))
        assert rel_err(got["asxtemp"], branch2_expected(
            "bgw" if sem == "bgw" else "gppker")) <= 1e-12
    for r in handle.batch([inp, inp], semantics=sem):
        check(r, ref, keys=("asxtemp",))
