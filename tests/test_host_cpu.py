"""Host-side checks that need no GPU: the Python binding's array validation."""
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
