"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
benchmark's CPU leg (baseline timing and result checks), as the checker.  The product path never
imports it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "liboracle.so"
REF_DIR = ORACLE_DIR / "_ref"

SEM = {"gppker": 0, "strict": 1, "bgw": 2}
ALL_BANDS_UNOCC = 1
SKIP_UNOCC = 2


class _Problem(C.Structure):
    _fields_ = [
        ("number_bands", C.c_int), ("nvband", C.c_int),
        ("ncouls", C.c_int), ("ngpown", C.c_int),
        ("aqsmtemp", C.c_void_p), ("aqsntemp", C.c_void_p),
        ("I_eps_array", C.c_void_p), ("wtilde_array", C.c_void_p),
        ("vcoul", C.c_void_p), ("inv_igp_index", C.c_void_p),
        ("indinv", C.c_void_p), ("wx_array", C.c_void_p),
        ("wx_per_band", C.c_int),
    ]


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise OSError(f"oracle not built: {LIB_PATH}")
        lib = C.CDLL(str(LIB_PATH))
        lib.gpp_oracle_sigma.argtypes = [C.POINTER(_Problem), C.c_int, C.c_int,
                                         C.c_void_p, C.c_void_p, C.c_void_p]
        lib.gpp_oracle_occ_sums.argtypes = [C.POINTER(_Problem), C.c_int,
                                            C.c_int, C.c_int, C.c_int,
                                            C.c_double, C.c_void_p, C.c_void_p]
        lib.gpp_oracle_resolve_igp.argtypes = [C.c_int, C.c_int, C.c_void_p,
                                               C.c_void_p]
        _lib = lib
    return _lib


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _problem(inp: dict):
    keep = {
        "aqsmtemp": _c(inp["aqsmtemp"], np.complex128),
        "aqsntemp": _c(inp["aqsntemp"], np.complex128),
        "I_eps_array": _c(inp["I_eps_array"], np.complex128),
        "wtilde_array": _c(inp["wtilde_array"], np.complex128),
        "vcoul": _c(inp["vcoul"], np.float64),
        "inv_igp_index": _c(inp["inv_igp_index"], np.int32),
        "indinv": _c(inp["indinv"], np.int32),
        "wx_array": _c(inp["wx_array"], np.float64),
    }
    nb = int(inp.get("number_bands", keep["aqsmtemp"].shape[0]))
    nc = int(inp.get("ncouls", keep["vcoul"].shape[0]))
    ng = int(inp.get("ngpown", keep["I_eps_array"].shape[0]))
    wx = keep["wx_array"]
    per_band = int(wx.ndim == 2 or wx.shape[0] != 3)
    p = _Problem(nb, int(inp["nvband"]), nc, ng,
                 *[keep[k].ctypes.data for k in (
                     "aqsmtemp", "aqsntemp", "I_eps_array", "wtilde_array",
                     "vcoul", "inv_igp_index", "indinv", "wx_array")],
                 per_band)
    return p, keep


def sigma(inp: dict, semantics="gppker", flags=0, want_achstemp=True,
          want_acht_n1=False) -> dict:
    """FP64 oracle of the full sum.  Returns achtemp[3], asxtemp[3]
    (complex), achstemp and optionally acht_n1_loc[NB]."""
    lib = load()
    p, keep = _problem(inp)
    out = np.zeros(12)
    ach = np.zeros(2) if want_achstemp else None
    n1 = np.zeros(2 * max(p.number_bands, 1)) if want_acht_n1 else None
    rc = lib.gpp_oracle_sigma(C.byref(p), SEM.get(semantics, semantics), flags,
                              out.ctypes.data,
                              ach.ctypes.data if ach is not None else None,
                              n1.ctypes.data if n1 is not None else None)
    if rc != 0:
        raise ValueError(f"oracle rejected the problem ({rc})")
    res = {
        "achtemp": out[0:6:2] + 1j * out[1:6:2],
        "asxtemp": out[6:12:2] + 1j * out[7:12:2],
    }
    if ach is not None:
        res["achstemp"] = complex(ach[0], ach[1])
    if n1 is not None:
        res["acht_n1_loc"] = n1[0::2][: p.number_bands] + 1j * n1[1::2][: p.number_bands]
    return res


def occ_sums(inp: dict, semantics: str, n1: int, my_igp: int, igp: int,
             wxt: float):
    """ssxt, scht of one (n1, my_igp, wxt) occupied row, unweighted."""
    lib = load()
    p, keep = _problem(inp)
    a = np.zeros(2)
    b = np.zeros(2)
    rc = lib.gpp_oracle_occ_sums(C.byref(p), SEM[semantics], n1, my_igp, igp,
                                 wxt, a.ctypes.data, b.ctypes.data)
    if rc != 0:
        raise ValueError("oracle rejected the row")
    return complex(a[0], a[1]), complex(b[0], b[1])


def resolve_igp(inp: dict) -> np.ndarray:
    lib = load()
    inv = _c(inp["inv_igp_index"], np.int32)
    ind = _c(inp["indinv"], np.int32)
    nc = int(inp.get("ncouls", ind.shape[0]))
    return np.array([lib.gpp_oracle_resolve_igp(i, nc, inv.ctypes.data,
                                                ind.ctypes.data)
                     for i in range(inv.shape[0])], np.int32)


# ---------------------------------------------------------------------------
# Full-frequency (FF) kernels: ff_oracle.c.

FF_SEM = {"cpp": 0, "f90": 1}
FF_OUTPUTS = ("asxDtemp", "achDtemp", "achDtemp_cor", "achDtemp_corb", "ach2Dtemp")


class _FFProblem(C.Structure):
    _fields_ = [
        ("number_bands", C.c_int), ("nvband", C.c_int), ("ncouls", C.c_int),
        ("ngpown", C.c_int), ("nFreq", C.c_int), ("nfreqeval", C.c_int),
        ("aqsmtemp", C.c_void_p), ("aqsntemp", C.c_void_p),
        ("I_epsR_array", C.c_void_p), ("I_epsA_array", C.c_void_p),
        ("vcoul", C.c_void_p), ("inv_igp_index", C.c_void_p),
        ("indinv", C.c_void_p), ("ekq", C.c_void_p), ("dFreqGrid", C.c_void_p),
        ("pref", C.c_void_p), ("dFreqBrd", C.c_void_p),
        ("schDt_matrix", C.c_void_p),
        ("freqevalmin", C.c_double), ("freqevalstep", C.c_double),
        ("pref_zb", C.c_double), ("occ", C.c_double),
    ]


_FF_ARRAYS = {
    "aqsmtemp": np.complex128, "aqsntemp": np.complex128,
    "I_epsR_array": np.complex128, "I_epsA_array": np.complex128,
    "vcoul": np.float64, "inv_igp_index": np.int32, "indinv": np.int32,
    "ekq": np.float64, "dFreqGrid": np.float64, "pref": np.float64,
    "dFreqBrd": np.complex128, "schDt_matrix": np.complex128,
}


def _ff_problem(inp: dict):
    keep = {k: _c(inp[k], t) for k, t in _FF_ARRAYS.items()}
    p = _FFProblem(int(inp["number_bands"]), int(inp["nvband"]),
                   int(inp["ncouls"]), int(inp["ngpown"]), int(inp["nFreq"]),
                   int(inp["nfreqeval"]),
                   *[keep[k].ctypes.data for k in _FF_ARRAYS],
                   float(inp["freqevalmin"]), float(inp["freqevalstep"]),
                   float(inp["pref_zb"]), float(inp["occ"]))
    return p, keep


def _ff_load():
    lib = load()
    if not hasattr(lib, "_ff_ready"):
        lib.ff_oracle.argtypes = [C.c_int, C.POINTER(_FFProblem)] + [C.c_void_p] * 6
        lib.ff_oracle.restype = C.c_int
        lib.ff_oracle_fact.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double,
                                       C.c_int, C.c_void_p, C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]
        lib.ff_oracle_fact.restype = C.c_int
        lib._ff_ready = True
    return lib


def ff_sigma(inp: dict, semantics: str = "cpp") -> dict:
    """All FF outputs: achsDtemp (complex) and the [nfreqeval] arrays."""
    lib = _ff_load()
    p, keep = _ff_problem(inp)
    ne = int(inp["nfreqeval"])
    achs = np.zeros(1, np.complex128)
    outs = {k: np.zeros(ne, np.complex128) for k in FF_OUTPUTS}
    rc = lib.ff_oracle(FF_SEM[semantics], C.byref(p), achs.ctypes.data,
                       *[outs[k].ctypes.data for k in FF_OUTPUTS])
    if rc:
        raise ValueError(f"ff_oracle rc={rc}")
    outs["achsDtemp"] = complex(achs[0])
    return outs


def ff_fact(semantics: str, loop: int, flag_occ: bool, wx: float, grid) -> tuple:
    """(ifreq, fact1, fact2) of the variant's frequency search."""
    lib = _ff_load()
    g = _c(grid, np.float64)
    f1, f2 = C.c_double(), C.c_double()
    f = lib.ff_oracle_fact(FF_SEM[semantics], loop, int(flag_occ), float(wx),
                           len(g), g.ctypes.data, C.byref(f1), C.byref(f2))
    return f, f1.value, f2.value
