"""Host-side checks that need no GPU: the Python binding's array validation,
the allocation-free entry point's refusals (they happen before its first HIP
call), and the per-thread error text of the handle-less entry points."""
import ctypes as C
import threading

import numpy as np
import pytest

from conftest import gppk


def _small(nb=4, nc=64, npg=8, nv=1, config=2):
    return gppk.fill_inputs(config, nb, nc, npg, nv, seed=3)


def test_make_problem_accepts_reference_dtypes():
    inp = _small()
    p = gppk.make_problem(inp)
    assert (p.number_bands, p.ncouls, p.ngpown, p.nvband) == (4, 64, 8, 1)
    # float64 with a trailing re/im axis is the same memory as complex128
    inp2 = dict(inp, aqsmtemp=inp["aqsmtemp"].view(np.float64).reshape(4, 64, 2))
    assert gppk.make_problem(inp2).aqsmtemp == inp["aqsmtemp"].ctypes.data


@pytest.mark.parametrize("key,bad,msg", [
    ("indinv", lambda a: a.astype(np.int64), "dtype int64"),
    ("inv_igp_index", lambda a: a.astype(np.float64), "dtype float64"),
    ("aqsmtemp", lambda a: a.astype(np.complex64), "dtype complex64"),
    ("vcoul", lambda a: a.astype(np.float32), "dtype float32"),
    ("I_eps_array", lambda a: a[:-1].copy(), "elements"),
    ("wtilde_array", lambda a: a[:, :10].copy(), "elements"),
    ("aqsntemp", lambda a: a[:2].copy(), "elements"),
    ("wx_array", lambda a: a[:2].copy(), "elements"),
    ("vcoul", lambda a: a[:5].copy(), "elements"),
])
def test_make_problem_rejects_wrong_dtype_or_size(key, bad, msg):
    inp = _small()
    inp[key] = bad(inp[key])
    with pytest.raises(ValueError, match=msg):
        gppk.make_problem(inp)


def test_make_problem_rejects_band_resolved_wx_too_short():
    inp = _small(config=3)
    assert inp["wx_array"].shape == (4, 3)
    inp["wx_array"] = inp["wx_array"][:3].copy()   # 9 values for 4 bands
    with pytest.raises(ValueError, match="wx_array holds 9"):
        gppk.make_problem(inp)


class _FakeDeviceArray:
    """Stands in for a GPU tensor: only the attributes the binding reads."""

    def __init__(self, like):
        self.dtype = like.dtype
        self.shape = like.shape
        self.ndim = like.ndim
        self.is_cuda = True
        self._n = like.size

    def numel(self):
        return self._n

    def is_contiguous(self):
        return True

    def data_ptr(self):
        return 0x1000


def test_make_problem_rejects_mixed_memory_spaces():
    inp = _small()
    inp["wtilde_array"] = _FakeDeviceArray(inp["wtilde_array"])
    with pytest.raises(ValueError, match="mix host and device"):
        gppk.make_problem(inp)


def test_make_ff_problem_checks_arrays():
    ff = gppk.ff_fill_inputs(2, 5, 2, 96, 8, 7, 9, semantics="f90", seed=1)
    gppk.make_ff_problem(ff)
    bad = dict(ff, indinv=ff["indinv"].astype(np.int64))
    with pytest.raises(ValueError, match="indinv: dtype int64"):
        gppk.make_ff_problem(bad)
    bad = dict(ff, I_epsR_array=ff["I_epsR_array"][:1].copy())
    with pytest.raises(ValueError, match="I_epsR_array holds"):
        gppk.make_ff_problem(bad)


# ---- gpp_sigma_noalloc: refusals that happen before the first HIP call ----

def _fake_problem(nb, nc, ng, nv=2):
    """A problem whose pointers are never dereferenced (the calls below are
    refused on the host)."""
    fake = C.c_void_p(0x1000)
    return gppk.Problem(nb, nv, nc, ng, *([fake] * 8), 0)


def _opts(lib, **kw):
    o = gppk.Options()
    lib.gpp_default_options(C.byref(o))
    o.memory = gppk.MEM_DEVICE
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _last_error(lib):
    return lib.gpp_last_error(None).decode()


def test_noalloc_refuses_host_memory_mixed_and_flags():
    lib = gppk.load_library()
    p = _fake_problem(8, 256, 32)
    out = C.c_void_p(0x2000)
    rc = lib.gpp_sigma_noalloc(C.byref(p), C.byref(_opts(lib, memory=gppk.MEM_HOST)), out)
    assert rc == gppk.GPP_EINVAL and "device-resident" in _last_error(lib)
    rc = lib.gpp_sigma_noalloc(C.byref(p), C.byref(_opts(lib, precision=gppk.PREC_MIXED)), out)
    assert rc == gppk.GPP_EUNSUPPORTED and "FP32 copy" in _last_error(lib)
    for flag in (gppk.FLAG_AUTOTUNE, gppk.FLAG_CHECK_FINITE):
        rc = lib.gpp_sigma_noalloc(C.byref(p), C.byref(_opts(lib, flags=flag)), out)
        assert rc == gppk.GPP_EUNSUPPORTED
    rc = lib.gpp_sigma_noalloc(C.byref(p), C.byref(_opts(lib)), None)
    assert rc == gppk.GPP_EINVAL and "null output" in _last_error(lib)
    rc = lib.gpp_sigma_noalloc(None, C.byref(_opts(lib)), out)
    assert rc == gppk.GPP_EINVAL


def _scratch_bytes(nb, nc, ng, nv, k=8, w=4, factored=False):
    """The entry point's scratch layout, recomputed: 256-byte slices for the
    counter, igp table, wx table (10 doubles per band), gathered aqsm and the
    block records."""
    def up(b):
        return (b + 255) // 256 * 256
    ngrp = (ng + 63) // 64
    n_un = ngrp * ((nc + 63) // 64) if factored else ngrp * ((nc + k * w - 1) // (k * w))
    n_occ = ng * ((nc + 2047) // 2048) if nv > 0 else 0
    n_ach = ng
    return (up(256) + up(ng * 4) + up(nb * 80) + up(nb * ng * 16)
            + up((n_un + n_occ + n_ach) * 96) + up(2 * 96))


@pytest.mark.parametrize("shape,fits", [
    ((512, 32768, 1638), True),      # canonical 512/2/32768/20: 19 MB
    ((1024, 32768, 1638), True),     # largest instance of the 64-problem batch
    ((2048, 65536, 3276), False),    # large 2048/16/65536/20: 118 MB
])
def test_noalloc_scratch_limit(shape, fits):
    lib = gppk.load_library()
    nb, nc, ng = shape
    need = _scratch_bytes(nb, nc, ng, 2)
    assert (need <= 64 << 20) == fits, need
    if fits:
        return  # an admitted problem would go on to the device
    p = _fake_problem(nb, nc, ng)
    rc = lib.gpp_sigma_noalloc(C.byref(p), C.byref(_opts(lib)), C.c_void_p(0x2000))
    assert rc == gppk.GPP_ESCRATCH
    err = _last_error(lib)
    assert f"needs {need} bytes" in err and str(64 << 20) in err, err


def test_handle_less_errors_are_per_thread():
    """Two threads fail differently at the same time; each reads back its
    own message from gpp_last_error(NULL)."""
    lib = gppk.load_library()
    start = threading.Barrier(2)
    seen = {}

    def worker(name, opts):
        p = _fake_problem(8, 256, 32)
        start.wait()
        for _ in range(200):
            rc = lib.gpp_sigma_noalloc(C.byref(p), C.byref(opts), C.c_void_p(0x2000))
            seen.setdefault(name, set()).add((rc, _last_error(lib)))

    t1 = threading.Thread(target=worker, args=("host", _opts(lib, memory=gppk.MEM_HOST)))
    t2 = threading.Thread(target=worker, args=("mixed", _opts(lib, precision=gppk.PREC_MIXED)))
    t1.start(); t2.start(); t1.join(); t2.join()
    assert len(seen["host"]) == 1 and len(seen["mixed"]) == 1, seen
    (rc1, e1), = seen["host"]
    (rc2, e2), = seen["mixed"]
    assert rc1 == gppk.GPP_EINVAL and "device-resident" in e1
    assert rc2 == gppk.GPP_EUNSUPPORTED and "FP32 copy" in e2
