"""Failure handling forced from the host: a simulated out-of-memory error at
every buffer growth of a call (GPP_FAIL_ALLOC_AFTER=N fails the N-th growth
of a handle) comes back as GPP_ENOMEM with its message, leaves the handle
usable, and leaks nothing; a failed host allocation (GPP_FAIL_HOST_ALLOC)
is turned into GPP_ENOMEM at the entry point.  The reference leaves its
allocations unchecked (gppKer.cpp:209-232)."""
import os

import numpy as np
import pytest

from conftest import gppk, gpp_oracle as orc, rel_err

pytestmark = pytest.mark.gpu


def test_simulated_allocation_failure_every_growth():
    import torch
    inp = gppk.fill_inputs(3, 8, 1024, 8, nvband=2, seed=9)
    ref = orc.sigma(inp, want_acht_n1=True)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    failed_at = []
    try:
        for n in range(1, 12):
            os.environ["GPP_FAIL_ALLOC_AFTER"] = str(n)
            try:
                h = gppk.Handle(0)
            except gppk.GppError as e:
                assert e.code == gppk.GPP_ENOMEM, e
                failed_at.append((n, "create"))
                continue
            try:
                h.sigma(inp, precision="mixed", acht_n1_loc=True)
                ok = True
            except gppk.GppError as e:
                assert e.code == gppk.GPP_ENOMEM, e
                assert "simulated allocation failure" in str(e)
                failed_at.append((n, "sigma"))
                ok = False
            del os.environ["GPP_FAIL_ALLOC_AFTER"]
            # the handle still works, failed or not
            got = h.sigma(inp, acht_n1_loc=True)
            for k in ("achtemp", "asxtemp", "acht_n1_loc"):
                assert rel_err(got[k], ref[k]) <= 1e-11, (n, k)
            h.close()
            if ok:
                break
    finally:
        os.environ.pop("GPP_FAIL_ALLOC_AFTER", None)
    assert failed_at and failed_at[0] == (1, "create")
    assert any(w == "sigma" for _, w in failed_at)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    # every handle freed its buffers (2 MiB growth granularity, runtime noise)
    assert abs(free1 - free0) <= (16 << 20), (free0, free1)


def test_host_allocation_failure_does_not_cross_the_c_abi():
    """A failed host allocation inside the batch and FF planning (forced
    with GPP_FAIL_HOST_ALLOC) throws std::bad_alloc inside the library; the
    entry point's handler returns GPP_ENOMEM with its message, nothing
    crosses the C ABI, and the same handle then computes the right sums."""
    insts = [gppk.fill_inputs(3, 6 + i, 320, 8, nvband=2, seed=70 + i) for i in range(3)]
    refs = [orc.sigma(d) for d in insts]
    ff_inp = gppk.ff_fill_inputs(2, 7, 3, 64, 4, 6, 10, semantics="f90", seed=3)
    ff_ref = orc.ff_sigma(ff_inp, "f90")
    h = gppk.Handle(0)
    try:
        os.environ["GPP_FAIL_HOST_ALLOC"] = "1"
        for call in (lambda: h.batch(insts), lambda: h.ff(ff_inp, semantics="f90")):
            with pytest.raises(gppk.GppError) as e:
                call()
            assert e.value.code == gppk.GPP_ENOMEM, e.value
            assert "host memory allocation failed" in str(e.value)
    finally:
        os.environ.pop("GPP_FAIL_HOST_ALLOC", None)
    for got, ref in zip(h.batch(insts), refs):
        for k in ("achtemp", "asxtemp"):
            assert rel_err(got[k], ref[k]) <= 1e-11, k
    got = h.ff(ff_inp, semantics="f90")
    for k in ("asxDtemp", "achDtemp", "achDtemp_cor", "ach2Dtemp"):
        scale = float(np.max(np.abs(ff_ref[k])))
        assert float(np.max(np.abs(got[k] - ff_ref[k]))) <= 1e-11 * scale, k
    h.close()
