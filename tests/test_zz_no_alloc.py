"""gpp_sigma_noalloc (the allocation-free entry point) against gpp_sigma_ex
on the same device-resident inputs, through the C ABI: two calls in a row,
the second on a smaller problem, in FP64 per term and band sum first, and
against the oracle."""
import ctypes as C

import numpy as np
import pytest

from conftest import gppk, gpp_oracle, gpu_available, rel_err

pytestmark = pytest.mark.gpu


def _noalloc(lib, inp, precision, semantics, tile_ig=8, waves=4, flags=0):
    import torch
    prob = gppk.make_problem(inp)
    opt = gppk.Options()
    lib.gpp_default_options(C.byref(opt))
    opt.precision = gppk.PRECISION[precision]
    opt.semantics = gppk.SEMANTICS[semantics]
    opt.memory = gppk.MEM_DEVICE
    opt.flags = flags
    opt.tile_ig = tile_ig
    opt.waves_per_block = waves
    opt.stream = torch.cuda.current_stream().cuda_stream
    out = torch.full((7,), complex("nan+nanj"), dtype=torch.complex128, device="cuda")
    rc = lib.gpp_sigma_noalloc(C.byref(prob), C.byref(opt), C.c_void_p(out.data_ptr()))
    assert rc == gppk.GPP_OK, (rc, lib.gpp_last_error(None).decode())
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    return {"achtemp": o[0:3], "asxtemp": o[3:6], "achstemp": o[6]}


SHAPES = [  # (config, NB, NV, NC, nodes_per_group): the second is smaller
    (2, 40, 2, 4096, 16),
    (3, 13, 3, 1500, 12),
]


@pytest.mark.parametrize("precision,semantics", [
    ("fp64", "gppker"), ("fp64", "bgw"), ("fp64", "strict"),
    ("fp64_factored", "gppker"), ("fp64_factored", "bgw"),
])
def test_noalloc_matches_sigma_ex_twice(handle, precision, semantics):
    lib = gppk.load_library()
    for cfg, nb, nv, nc, npg in SHAPES:
        if precision == "fp64_factored" and cfg == 3:
            cfg = 2  # the band sum first needs shared frequencies
        print(f"noalloc {precision} {semantics} config {cfg} {nb}/{nv}/{nc}/{npg}",
              flush=True)
        inp = gppk.fill_inputs_device(cfg, nb, nc, npg, nv, seed=11 + nb)
        got = _noalloc(lib, inp, precision, semantics)
        ref = handle.sigma(inp, precision=precision, semantics=semantics,
                           tile_ig=8, waves_per_block=4)
        for k in ("achtemp", "asxtemp"):
            assert np.array_equal(got[k], ref[k]), (k, got[k], ref[k])
        assert got["achstemp"] == ref["achstemp"]


def test_noalloc_matches_oracle_and_flags(handle):
    lib = gppk.load_library()
    cfg, nb, nv, nc, npg = 3, 12, 2, 900, 9
    host = gppk.fill_inputs(cfg, nb, nc, npg, nv, seed=5)
    import torch
    dev = {k: (torch.from_numpy(v).cuda() if isinstance(v, np.ndarray) else v)
           for k, v in host.items()}
    print("noalloc vs oracle 12/2/900/9", flush=True)
    got = _noalloc(lib, dev, "fp64", "gppker", tile_ig=0, waves=0)
    ref = gpp_oracle.sigma(host)
    assert rel_err(got["achtemp"], ref["achtemp"]) <= 1e-11
    assert rel_err(got["asxtemp"], ref["asxtemp"]) <= 1e-11
    # band-range flags and NO_ACHSTEMP as gpp_sigma_ex gives them
    for flags in (gppk.FLAG_SKIP_UNOCC, gppk.FLAG_ALL_BANDS_UNOCC | gppk.FLAG_NO_ACHSTEMP):
        got = _noalloc(lib, dev, "fp64", "gppker", flags=flags)
        exp = handle.sigma(dev, precision="fp64", flags=flags, tile_ig=8,
                           waves_per_block=4)
        assert np.array_equal(got["achtemp"], exp["achtemp"])
        assert np.array_equal(got["asxtemp"], exp["asxtemp"])
        assert got["achstemp"] == exp["achstemp"]
