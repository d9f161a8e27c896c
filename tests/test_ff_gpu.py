"""GPU parity tests of the full-frequency (FF) path through the C ABI
(include/ff_sigma.h) against the CPU oracle and the reference's own
outputs.  FP64: within 1e-11 of the largest entry of each output array."""
import ctypes
import json

import numpy as np
import pytest

from conftest import GOLDEN, gppk, gpp_oracle as orc
from test_ff_cpu import (GRID_CASE_NAMES, KATS, LINEAR_EPS, SHAPES, c,
                         closed_form, grid_case, kat_problem, linear_eps, rel)

pytestmark = pytest.mark.gpu

TOL = 1e-11
OUTS = ("asxDtemp", "achDtemp", "achDtemp_cor", "achDtemp_corb", "ach2Dtemp")


def say(*a):
    print(*a, flush=True)


def check(got, ref, keys=OUTS, tol=TOL):
    assert abs(got["achsDtemp"] - ref["achsDtemp"]) <= tol * max(abs(ref["achsDtemp"]), 1e-300), \
        (got["achsDtemp"], ref["achsDtemp"])
    for k in keys:
        e = rel(got[k], ref[k])
        assert e <= tol, (k, e, got[k][:3], ref[k][:3])


@pytest.mark.parametrize("shape", [(5, 2, 96, 8, 7, 9), (20, 6, 130, 11, 9, 13),
                                   (3, 3, 300, 2, 4, 40), (9, 0, 70, 7, 5, 6)])
@pytest.mark.parametrize("sem", ["cpp", "f90"])
def test_random_vs_oracle(handle, shape, sem):
    """Seeded random inputs (non-uniform grid, band energies on both sides of
    the grid); 20 bands put more than 8 bands on one slice (several passes)."""
    say("ff random", shape, sem)
    inp = gppk.ff_fill_inputs(2, *shape, semantics=sem, seed=sum(shape))
    got = handle.ff(inp, semantics=sem)
    check(got, orc.ff_sigma(inp, sem))


@pytest.mark.parametrize("shape", [(1, 1, 1, 1, 2, 1), (2, 0, 5, 3, 2, 1),
                                   (3, 3, 17, 4, 3, 2), (9, 1, 33, 33, 2, 3)])
@pytest.mark.parametrize("sem", ["cpp", "f90"])
def test_edge_shapes_vs_oracle(handle, shape, sem):
    """Single band / column / row, two grid points, one evaluation
    frequency, nvband = 0 and nvband = number_bands, NG = NC."""
    say("ff edge", shape, sem)
    inp = gppk.ff_fill_inputs(2, *shape, semantics=sem, seed=3 + sum(shape))
    check(handle.ff(inp, semantics=sem), orc.ff_sigma(inp, sem))


@pytest.mark.parametrize("sem", ["cpp", "f90"])
def test_unsorted_grid_vs_oracle(handle, sem):
    """A frequency grid out of order: the plan falls back from the monotone
    walker to the reference's linear scans."""
    say("ff unsorted grid", sem)
    inp = gppk.ff_fill_inputs(2, 6, 2, 48, 5, 9, 11, semantics=sem, seed=21)
    perm = np.random.default_rng(5).permutation(9)
    inp = dict(inp, dFreqGrid=np.ascontiguousarray(inp["dFreqGrid"][perm]),
               pref=np.ascontiguousarray(inp["pref"][perm]),
               dFreqBrd=np.ascontiguousarray(inp["dFreqBrd"][perm]))
    assert np.any(np.diff(inp["dFreqGrid"]) < 0)
    check(handle.ff(inp, semantics=sem), orc.ff_sigma(inp, sem))


@pytest.mark.parametrize("r", KATS, ids=lambda r: f"seed{r['seed']}")
@pytest.mark.parametrize("sem", ["cpp", "f90"])
def test_kat_inputs_vs_reference_and_oracle(handle, r, sem):
    say("ff kat", r["seed"], sem)
    inp = kat_problem(r)
    got = handle.ff(inp, semantics=sem)
    check(got, orc.ff_sigma(inp, sem))
    if sem == "cpp":  # the reference's own kernel outputs
        assert abs(got["achsDtemp"] - complex(*r["achsDtemp"])) <= TOL * abs(complex(*r["achsDtemp"]))
        for k in ("asxDtemp", "achDtemp", "achDtemp_cor"):
            assert rel(got[k], c(r[k])) <= TOL, k


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("sem", ["cpp", "f90"])
def test_constant_fills_closed_form(handle, shape, sem):
    say("ff closed form", shape, sem)
    inp = gppk.ff_fill_inputs(1, *shape, semantics=sem)
    got = handle.ff(inp, semantics=sem)
    achs, asx, cor, a2 = closed_form(sem, *shape)
    assert abs(got["achsDtemp"] - achs) <= 1e-12 * abs(achs)
    assert rel(got["asxDtemp"], asx) <= 1e-12
    assert rel(got["achDtemp_cor"], cor) <= 1e-12
    assert rel(got["ach2Dtemp"], a2) <= 1e-12


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("sem", ["cpp", "f90"])
def test_linear_in_frequency_closed_form(handle, shape, sem):
    """I_eps linear in the grid frequency: the slice sums and the host's
    interpolation weights reproduce the linear function at |wx|."""
    say("ff linear eps", shape, sem)
    inp = linear_eps(gppk.ff_fill_inputs(1, *shape, semantics=sem), *LINEAR_EPS)
    got = handle.ff(inp, semantics=sem)
    achs, asx, cor, a2 = closed_form(sem, *shape, *LINEAR_EPS)
    assert abs(got["achsDtemp"] - achs) <= 1e-12 * abs(achs)
    assert rel(got["asxDtemp"], asx) <= 1e-12
    assert rel(got["achDtemp_cor"], cor) <= 1e-12
    assert rel(got["ach2Dtemp"], a2) <= 1e-12


@pytest.mark.parametrize("name", GRID_CASE_NAMES)
def test_f90_grid_terms_closed_forms(handle, name):
    """The frequency-grid kernel (ffKernel.f90 semantics) against the closed
    forms of single-point, constant and linear schDt_matrix."""
    name, inp, want = grid_case(name)
    say("ff grid closed form", name)
    got = handle.ff(inp, semantics="f90")
    for k, v in want.items():
        assert rel(got[k], v) <= 1e-12, (name, k)


def test_device_memory_and_generator(handle):
    """Device-resident inputs from the device generator: the same bits as the
    host generator, and the same results as host inputs."""
    import torch
    say("ff device")
    shape = (6, 2, 200, 10, 8, 12)
    host = gppk.ff_fill_inputs(2, *shape, semantics="f90", seed=77)
    dev = gppk.ff_fill_inputs_device(2, *shape, semantics="f90", seed=77)
    torch.cuda.synchronize()
    for k in gppk.FF_ARRAYS:
        assert np.array_equal(dev[k].cpu().numpy(), host[k]), k
    a = handle.ff(host, semantics="f90")
    b = handle.ff(dev, semantics="f90", memory=gppk.MEM_DEVICE,
                  stream=torch.cuda.current_stream().cuda_stream)
    for k in ("achsDtemp",) + OUTS:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


def test_parts_add_up(handle):
    inp = gppk.ff_fill_inputs(2, 7, 3, 64, 4, 6, 10, semantics="f90", seed=3)
    full = handle.ff(inp, semantics="f90")
    sep = {}
    for part, keys in ((gppk.FF_PART_ACHS, ()), (gppk.FF_PART_ASX, ("asxDtemp",)),
                       (gppk.FF_PART_ACH, ("achDtemp", "achDtemp_corb")),
                       (gppk.FF_PART_COR, ())):
        sep[part] = handle.ff(inp, semantics="f90", parts=part)
    assert sep[gppk.FF_PART_ACHS]["achsDtemp"] == full["achsDtemp"]
    assert np.array_equal(sep[gppk.FF_PART_ASX]["asxDtemp"], full["asxDtemp"])
    assert np.array_equal(sep[gppk.FF_PART_ACH]["achDtemp"], full["achDtemp"])
    for k in ("achDtemp_cor", "ach2Dtemp"):
        s = sep[gppk.FF_PART_ACH][k] + sep[gppk.FF_PART_COR][k]
        assert rel(s, full[k]) <= 1e-14, k


def test_dropin_kernels_match_reference(handle):
    """ff_*_Kernel called like the reference's functions (host pointers,
    adding into the caller's arrays) reproduce the reference's own outputs."""
    say("ff drop-ins")
    lib = gppk.load_library()
    r = KATS[0]
    d = kat_problem(r)
    P = lambda a: a.ctypes.data
    nb, nv, nc, ng, nf, ne = (r[k] for k in ("nb", "nv", "nc", "ng", "nf", "ne"))
    keep = {k: np.ascontiguousarray(v) for k, v in d.items() if isinstance(v, np.ndarray)}
    achs = np.zeros(1, np.complex128)
    assert lib.ff_achsDtemp_Kernel(nb, ng, nc, P(keep["inv_igp_index"]), P(keep["indinv"]),
                                   P(keep["aqsntemp"]), P(keep["aqsmtemp"]),
                                   P(keep["I_epsR_array"]), P(keep["vcoul"]), P(achs)) == 0
    assert abs(achs[0] - complex(*r["achsDtemp"])) <= TOL * abs(complex(*r["achsDtemp"]))
    asx = np.ones(ne, np.complex128)  # the kernel adds into it
    scratch = np.zeros(64, np.complex128)
    assert lib.ff_asxDtemp_Kernel(nv, ne, nc, ng, 1, nf, ctypes.c_double(r["fmin"]),
                                  ctypes.c_double(r["fstep"]), ctypes.c_double(r["occ"]),
                                  P(keep["ekq"]), P(keep["dFreqGrid"]),
                                  P(keep["inv_igp_index"]), P(keep["indinv"]),
                                  P(keep["aqsmtemp"]), P(keep["aqsntemp"]), P(keep["vcoul"]),
                                  P(keep["I_epsR_array"]), P(scratch), P(asx)) == 0
    assert rel(asx - 1, c(r["asxDtemp"])) <= TOL
    ach = np.zeros(ne, np.complex128)
    s1, s2, s3 = (np.zeros(ne, np.complex128) for _ in range(3))
    assert lib.ff_achDtemp_Kernel(nb, nv, ne, nc, ng, nf, ctypes.c_double(r["fmin"]),
                                  ctypes.c_double(r["fstep"]), P(keep["ekq"]),
                                  ctypes.c_double(r["pref_zb"]), P(keep["pref"]),
                                  P(keep["dFreqGrid"]), P(keep["dFreqBrd"]),
                                  P(keep["schDt_matrix"]), P(s1), P(s2), P(s3), P(ach)) == 0
    assert rel(ach, c(r["achDtemp"])) <= TOL
    cor, corb, a2 = (np.zeros(ne, np.complex128) for _ in range(3))
    assert lib.ff_achDtemp_cor_Kernel(nb, nv, ne, nc, ng, nf, ctypes.c_double(r["fmin"]),
                                      ctypes.c_double(r["fstep"]), P(keep["ekq"]),
                                      P(keep["dFreqGrid"]), P(keep["inv_igp_index"]),
                                      P(keep["indinv"]), P(keep["aqsmtemp"]),
                                      P(keep["aqsntemp"]), P(keep["vcoul"]),
                                      P(keep["I_epsR_array"]), P(keep["I_epsA_array"]),
                                      P(s1), P(s2), P(s3), P(a2), P(cor), P(corb)) == 0
    assert rel(cor, c(r["achDtemp_cor"])) <= TOL


def test_invalid_arguments(handle):
    inp = gppk.ff_fill_inputs(2, 3, 1, 32, 4, 3, 4, seed=1)
    with pytest.raises(gppk.GppError) as e:
        handle.ff(dict(inp, nFreq=1))
    assert e.value.code == gppk.GPP_EINVAL
    with pytest.raises(gppk.GppError):
        handle.ff(dict(inp, I_epsA_array=None))
    with pytest.raises(gppk.GppError):
        handle.ff(inp, parts=64)


def test_cli_matches_reference_binary_lines():
    """`ff` (device-generated mini-app fills) prints the reference binary's
    result lines."""
    import subprocess
    r = json.loads((GOLDEN / "ref_ff_constfill.json").read_text())[0]
    args = [str(r[k]) for k in ("number_bands", "nvband", "ncouls", "ngpown", "nFreq", "nfreqeval")]
    out = subprocess.run([str(gppk.FF_CLI_PATH)] + args, capture_output=True,
                         text=True, timeout=120).stdout
    lines = {l.split(" = ")[0]: l.split(" = ")[1] for l in out.splitlines() if " = (" in l}
    for k in ("achsDtemp", "asxDtemp", "achDtemp_cor"):
        want = "( {:f}, {:f}) ".format(*r[k])
        assert lines[k] == want, (k, lines[k], want)


@pytest.mark.parametrize("sem", ["cpp", "f90"])
def test_cli_lines_match_closed_form(sem):
    """`ff ... --semantics S` (device-generated mini-app fills) prints the
    closed-form achsDtemp, asxDtemp[0] and achDtemp_cor[0] (the C++ line of
    asxDtemp includes achDtemp, zero for these fills)."""
    import subprocess
    shape = (8, 2, 64, 8, 6, 5)
    out = subprocess.run([str(gppk.FF_CLI_PATH)] + [str(x) for x in shape] +
                         ["--semantics", sem], capture_output=True, text=True,
                         timeout=120, check=True).stdout
    lines = {l.split(" = ")[0]: l.split(" = ")[1] for l in out.splitlines() if " = (" in l}
    achs, asx, cor, _ = closed_form(sem, *shape)
    for k, v in (("achsDtemp", achs), ("asxDtemp", asx[0]), ("achDtemp_cor", cor[0])):
        assert lines[k] == "( {:f}, {:f}) ".format(v.real, v.imag), (k, lines[k], v)


def _ff_sweep_shapes(n, seed):
    """Deterministic ragged FF shapes (number_bands, nvband, ncouls, ngpown,
    nFreq, nfreqeval)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nb = int(rng.integers(1, 20))
        nv = int(rng.integers(0, nb + 1))
        nc = int(rng.integers(1, 700))
        ng = int(rng.integers(1, min(nc, 40) + 1))
        nf = int(rng.integers(2, 60))
        ne = int(rng.integers(1, 80))
        out.append((nb, nv, nc, ng, nf, ne))
    return out


@pytest.mark.parametrize("shape", _ff_sweep_shapes(16, 2025))
def test_shape_sweep_vs_oracle(handle, shape):
    """Ragged random shapes in both semantics (several passes per slice when
    more than 8 bands need it, splits of the frequency sum, odd grid-point
    counts for the paired reciprocals)."""
    say("ff sweep", shape)
    for sem in ("cpp", "f90"):
        inp = gppk.ff_fill_inputs(2, *shape, semantics=sem, seed=sum(shape) + 1)
        check(handle.ff(inp, semantics=sem), orc.ff_sigma(inp, sem))
