"""Python host binding of the MI355X GPP sigma library (lib/libgpp.so).

The compute path is the C ABI declared in include/gpp_sigma.h; this module
only wraps it with ctypes so that tests, the benchmark and Python callers
can hand it numpy arrays (host memory) or torch device tensors (device
memory).  It never computes sigma itself: if the shared library is missing
the loader raises, there is no fallback.

Reference interface mirrored: the gppKer loop nest (gppKer.cpp:270-361) and
the operator split tillNvband_kernel / achtemp_kernel
(gppKer_gpuComplex.cpp:127,148); for the full-frequency kernels
(include/ff_sigma.h) the kernels of FFKer_cori/ffKernel.cpp:107-300.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "lib" / "libgpp.so"
CLI_PATH = PKG_DIR / "bin" / "gpp"
FF_CLI_PATH = PKG_DIR / "bin" / "ff"

GPP_OK = 0
GPP_EINVAL = -1
GPP_ENOMEM = -2
GPP_EDEVICE = -3
GPP_ENONFINITE = -4
GPP_EUNSUPPORTED = -5
GPP_ESCRATCH = -6

PREC_FP64 = 0
PREC_MIXED = 1
PREC_FP64_FACTORED = 2
SEM_GPPKER = 0
SEM_GPPKER_STRICT = 1
SEM_BGW = 2
MEM_HOST = 0
MEM_DEVICE = 1
FLAG_ALL_BANDS_UNOCC = 1
FLAG_SKIP_UNOCC = 2
FLAG_CHECK_FINITE = 4
FLAG_NO_ACHSTEMP = 8
FLAG_AUTOTUNE = 16
FLAG_MIXED_KAHAN = 32

SEMANTICS = {"gppker": SEM_GPPKER, "strict": SEM_GPPKER_STRICT, "bgw": SEM_BGW}
PRECISION = {"fp64": PREC_FP64, "mixed": PREC_MIXED,
             "fp64_factored": PREC_FP64_FACTORED}

FF_SEM_CPP = 0
FF_SEM_F90 = 1
FF_SEMANTICS = {"cpp": FF_SEM_CPP, "f90": FF_SEM_F90}
FF_PART_ACHS = 1
FF_PART_ASX = 2
FF_PART_ACH = 4
FF_PART_COR = 8
FF_OUTPUTS = ("asxDtemp", "achDtemp", "achDtemp_cor", "achDtemp_corb", "ach2Dtemp")

# Exported symbols, one per declaration of include/gpp_sigma.h and
# include/ff_sigma.h.
EXPORTS = (
    "gpp_create", "gpp_destroy", "gpp_last_error", "gpp_default_options",
    "gpp_sigma_ex", "gpp_sigma_batch", "gpp_autotune", "gpp_sigma",
    "gpp_achtemp_kernel", "gpp_tillNvband_kernel", "gpp_fill_inputs",
    "gpp_fill_inputs_device", "gpp_sigma_noalloc", "gpp_version",
    "ff_default_options", "ff_sigma_ex", "ff_achsDtemp_Kernel",
    "ff_asxDtemp_Kernel", "ff_achDtemp_Kernel", "ff_achDtemp_cor_Kernel",
    "ff_fill_inputs", "ff_fill_inputs_device",
)


class c128(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class Problem(C.Structure):
    _fields_ = [
        ("number_bands", C.c_int), ("nvband", C.c_int),
        ("ncouls", C.c_int), ("ngpown", C.c_int),
        ("aqsmtemp", C.c_void_p), ("aqsntemp", C.c_void_p),
        ("I_eps_array", C.c_void_p), ("wtilde_array", C.c_void_p),
        ("vcoul", C.c_void_p), ("inv_igp_index", C.c_void_p),
        ("indinv", C.c_void_p), ("wx_array", C.c_void_p),
        ("wx_per_band", C.c_int),
    ]


class Options(C.Structure):
    _fields_ = [
        ("precision", C.c_int), ("semantics", C.c_int),
        ("memory", C.c_int), ("flags", C.c_int),
        ("stream", C.c_void_p), ("resonance_threshold", C.c_double),
        ("tile_ig", C.c_int), ("waves_per_block", C.c_int),
    ]


class Result(C.Structure):
    _fields_ = [
        ("achtemp", c128 * 3), ("asxtemp", c128 * 3), ("achstemp", c128),
        ("acht_n1_loc", C.c_void_p),
        ("unocc_kernel_ms", C.c_double), ("occ_kernel_ms", C.c_double),
        ("total_ms", C.c_double), ("terms", C.c_double),
        ("fp64_fallback_terms", C.c_longlong),
    ]


_lib = None


class FFProblem(C.Structure):
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


class FFOptions(C.Structure):
    _fields_ = [("semantics", C.c_int), ("memory", C.c_int),
                ("stream", C.c_void_p), ("parts", C.c_int)]


class FFResult(C.Structure):
    _fields_ = [
        ("achsDtemp", c128), ("asxDtemp", C.c_void_p), ("achDtemp", C.c_void_p),
        ("achDtemp_cor", C.c_void_p), ("achDtemp_corb", C.c_void_p),
        ("ach2Dtemp", C.c_void_p), ("slice_kernel_ms", C.c_double),
        ("freq_kernel_ms", C.c_double), ("total_ms", C.c_double),
        ("slice_bytes", C.c_double), ("slices", C.c_longlong),
        ("slice_terms", C.c_longlong), ("freq_terms", C.c_longlong),
        ("host_plan_ms", C.c_double), ("host_assemble_ms", C.c_double),
    ]


FF_ARRAYS = ("aqsmtemp", "aqsntemp", "I_epsR_array", "I_epsA_array", "vcoul",
             "inv_igp_index", "indinv", "ekq", "dFreqGrid", "pref", "dFreqBrd",
             "schDt_matrix")


def make_ff_problem(inp: dict) -> FFProblem:
    """FFProblem from a dict of arrays (numpy host or torch device) with the
    keys of FF_ARRAYS plus number_bands, nvband, ncouls, ngpown, nFreq,
    nfreqeval, freqevalmin, freqevalstep, pref_zb, occ.  Arrays are checked
    as in make_problem (dtype, element count, one memory space); absent
    optional ones pass NULL."""
    nb, nc, ng = int(inp["number_bands"]), int(inp["ncouls"]), int(inp["ngpown"])
    nf = int(inp["nFreq"])
    ptrs = _checked(inp, (
        ("aqsmtemp", "c128", nb * nc, True),
        ("aqsntemp", "c128", nb * nc, True),
        ("I_epsR_array", "c128", nf * ng * nc, False),
        ("I_epsA_array", "c128", nf * ng * nc, False),
        ("vcoul", "f64", nc, True),
        ("inv_igp_index", "i32", ng, True),
        ("indinv", "i32", nc, True),
        ("ekq", "f64", nb, False),
        ("dFreqGrid", "f64", nf, False),
        ("pref", "f64", nf, False),
        ("dFreqBrd", "c128", nf, False),
        ("schDt_matrix", "c128", nb * nf, False),
    ), "ff problem")
    return FFProblem(nb, int(inp["nvband"]), nc, ng, nf,
                     int(inp["nfreqeval"]), *ptrs,
                     float(inp["freqevalmin"]), float(inp["freqevalstep"]),
                     float(inp["pref_zb"]), float(inp["occ"]))


def load_library(path: os.PathLike | str | None = None) -> C.CDLL:
    """Load libgpp.so (built in-tree by __graft_entry__.build()).  Raises
    OSError when it is missing: there is no CPU fallback."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise OSError(f"GPP library not built: {p} (run __graft_entry__.build())")
    # PyTorch ships its own HIP runtime (soname libamdhip64.so.7, the same as
    # /opt/rocm's, which libgpp.so links).  Loaded first, it also serves this
    # library; loaded second, it would bring a second HIP runtime into the
    # process, and the two fail to share the device.  So torch, when
    # present, is imported before the library is opened.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(p))
    vp, ip = C.c_void_p, C.c_int
    lib.gpp_create.argtypes = [C.POINTER(vp), ip]
    lib.gpp_destroy.argtypes = [vp]
    lib.gpp_last_error.argtypes = [vp]
    lib.gpp_last_error.restype = C.c_char_p
    lib.gpp_default_options.argtypes = [C.POINTER(Options)]
    lib.gpp_sigma_ex.argtypes = [vp, C.POINTER(Problem), C.POINTER(Options),
                                 C.POINTER(Result)]
    lib.gpp_sigma_batch.argtypes = [vp, C.POINTER(Problem), ip,
                                    C.POINTER(Options), C.POINTER(Result)]
    lib.gpp_autotune.argtypes = [vp, C.POINTER(Problem), C.POINTER(Options),
                                 C.POINTER(ip), C.POINTER(ip)]
    lib.gpp_sigma.argtypes = [ip, ip, ip, ip] + [vp] * 10
    lib.gpp_achtemp_kernel.argtypes = [ip, ip] + [vp] * 8 + [ip] + [vp] * 6
    lib.gpp_tillNvband_kernel.argtypes = ([vp] * 8 + [ip, ip, ip, vp, ip, ip,
                                          C.c_double, ip])
    lib.gpp_fill_inputs.argtypes = [ip, C.c_ulonglong, ip, ip, ip, ip] + [vp] * 8
    lib.gpp_fill_inputs_device.argtypes = ([ip, C.c_ulonglong, ip, ip, ip, ip]
                                           + [vp] * 9)
    lib.gpp_sigma_noalloc.argtypes = [C.POINTER(Problem), C.POINTER(Options), vp]
    lib.gpp_version.restype = C.c_char_p
    lib.ff_default_options.argtypes = [C.POINTER(FFOptions)]
    lib.ff_sigma_ex.argtypes = [vp, C.POINTER(FFProblem), C.POINTER(FFOptions),
                                C.POINTER(FFResult)]
    lib.ff_achsDtemp_Kernel.argtypes = [ip, ip, ip] + [vp] * 7
    lib.ff_asxDtemp_Kernel.argtypes = ([ip] * 6 + [C.c_double] * 3 + [vp] * 10)
    lib.ff_achDtemp_Kernel.argtypes = ([ip] * 6 + [C.c_double] * 2 + [vp, C.c_double]
                                       + [vp] * 8)
    lib.ff_achDtemp_cor_Kernel.argtypes = ([ip] * 6 + [C.c_double] * 2 + [vp] * 15)
    fill = [ip, ip, C.c_ulonglong, ip, ip, ip, ip] + [vp] * 13
    lib.ff_fill_inputs.argtypes = fill
    lib.ff_fill_inputs_device.argtypes = fill + [vp]
    if path is None:
        _lib = lib
    return lib


class GppError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gpp error {code}: {msg}")
        self.code = code


def _ptr(a) -> int | None:
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("arrays must be C-contiguous")
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensors must be contiguous")
        return a.data_ptr()
    raise TypeError(f"unsupported array type {type(a)}")


# Element kinds the C ABI reads: complex128 (or float64 with a trailing
# re/im axis of 2), float64, int32.
_KIND_DTYPES = {
    "c128": ("complex128",),
    "f64": ("float64",),
    "i32": ("int32",),
}


def _space(a) -> str:
    """'host' for numpy arrays and CPU tensors, 'device' for GPU tensors."""
    if isinstance(a, np.ndarray):
        return "host"
    if hasattr(a, "data_ptr"):
        return "device" if getattr(a, "is_cuda", False) else "host"
    raise TypeError(f"unsupported array type {type(a)}")


def _elements(a, kind: str, name: str) -> int:
    """Element count of `a` as the C side reads it; raises ValueError when
    its dtype is not the one the C ABI expects for `kind`."""
    dt = str(a.dtype).replace("torch.", "")
    size = int(a.size) if isinstance(a, np.ndarray) else int(a.numel())
    if dt in _KIND_DTYPES[kind]:
        return size
    if kind == "c128" and dt == "float64" and a.ndim >= 1 and a.shape[-1] == 2:
        return size // 2
    raise ValueError(f"{name}: dtype {dt}, expected "
                     f"{'complex128' if kind == 'c128' else _KIND_DTYPES[kind][0]}")


def _checked(inp: dict, spec, what: str) -> list:
    """Checks every array of `spec` ((key, kind, minimum element count,
    required) tuples) against its dtype and size and that all of them live
    in one memory space; returns their pointers in order (None for absent
    optional arrays)."""
    spaces = set()
    ptrs = []
    for key, kind, need, required in spec:
        a = inp.get(key)
        if a is None:
            if required and need > 0:
                raise ValueError(f"{what}: {key} is missing")
            ptrs.append(None)
            continue
        n = _elements(a, kind, key)
        if n < need:
            raise ValueError(f"{what}: {key} holds {n} elements, the shape "
                             f"needs {need}")
        spaces.add(_space(a))
        ptrs.append(_ptr(a))
    if len(spaces) > 1:
        raise ValueError(f"{what}: arrays mix host and device memory")
    return ptrs


def make_problem(inp: dict) -> Problem:
    """Problem struct from a dict of arrays (numpy host or torch device).
    Keys: aqsmtemp, aqsntemp, I_eps_array, wtilde_array (complex128, or
    float64 with a trailing re/im axis), vcoul (float64), inv_igp_index,
    indinv (int32), wx_array (float64, [3] or [NB,3]), nvband, and
    optionally number_bands/ncouls/ngpown.  Every array is checked against
    the dtype and element count the shape needs (ValueError otherwise), and
    all must be in one memory space: the C side reads them as raw pointers."""
    aqsm = inp["aqsmtemp"]
    ieps = inp["I_eps_array"]
    nb = int(inp.get("number_bands", aqsm.shape[0]))
    nc = int(inp.get("ncouls", aqsm.shape[1] if aqsm.ndim >= 2 else inp["vcoul"].shape[0]))
    ng = int(inp.get("ngpown", ieps.shape[0]))
    wx = inp["wx_array"]
    per_band = int(wx.ndim == 2 or wx.shape[0] != 3)
    ptrs = _checked(inp, (
        ("aqsmtemp", "c128", nb * nc, True),
        ("aqsntemp", "c128", nb * nc, True),
        ("I_eps_array", "c128", ng * nc, True),
        ("wtilde_array", "c128", ng * nc, True),
        ("vcoul", "f64", nc, True),
        ("inv_igp_index", "i32", ng, True),
        ("indinv", "i32", nc, True),
        ("wx_array", "f64", 3 * nb if per_band else 3, True),
    ), "gpp problem")
    return Problem(nb, int(inp["nvband"]), nc, ng, *ptrs, per_band)


class BatchProblems:
    """The C problem table of a batch, built once from a list of input
    dicts.  Passed to Handle.batch in place of the list it skips rebuilding
    the table on every call; the arrays must stay alive and in place (their
    contents may change between calls)."""

    def __init__(self, inputs: list):
        self.inputs = list(inputs)
        self.table = (Problem * len(self.inputs))(
            *[make_problem(d) for d in self.inputs])

    def __len__(self) -> int:
        return len(self.inputs)


# numpy view of gpp_result (include/gpp_sigma.h), for converting a batch's
# results in bulk.
_RESULT_DTYPE = np.dtype([
    ("achtemp", np.complex128, (3,)), ("asxtemp", np.complex128, (3,)),
    ("achstemp", np.complex128), ("acht_n1_loc", np.uintp),
    ("unocc_kernel_ms", np.float64), ("occ_kernel_ms", np.float64),
    ("total_ms", np.float64), ("terms", np.float64),
    ("fp64_fallback_terms", np.int64)])


def _batch_results(res, n: int) -> list:
    if _RESULT_DTYPE.itemsize != C.sizeof(Result):
        raise RuntimeError("gpp_result layout mismatch")
    a = np.frombuffer(res, dtype=_RESULT_DTYPE, count=n)
    ach, asx = a["achtemp"].copy(), a["asxtemp"].copy()
    achs = a["achstemp"].tolist()
    rest = [a[k].tolist() for k in ("unocc_kernel_ms", "occ_kernel_ms",
                                    "total_ms", "terms", "fp64_fallback_terms")]
    return [{"achtemp": ach[i], "asxtemp": asx[i], "achstemp": achs[i],
             "unocc_kernel_ms": rest[0][i], "occ_kernel_ms": rest[1][i],
             "total_ms": rest[2][i], "terms": rest[3][i],
             "fp64_fallback_terms": rest[4][i]} for i in range(n)]


def _cplx(z: c128) -> complex:
    return complex(z.re, z.im)


class Handle:
    """Owns one gpp_handle (device buffers, stream, events)."""

    def __init__(self, device: int = -1, lib: C.CDLL | None = None):
        self.lib = lib or load_library()
        h = C.c_void_p()
        rc = self.lib.gpp_create(C.byref(h), device)
        if rc != GPP_OK:
            raise GppError(rc, "gpp_create failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.gpp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def last_error(self) -> str:
        return self.lib.gpp_last_error(self.h).decode()

    def _options(self, inp, precision, semantics, memory, flags, stream,
                 resonance_threshold, tile_ig, waves_per_block) -> Options:
        if memory is None:
            memory = MEM_DEVICE if hasattr(inp["aqsmtemp"], "data_ptr") else MEM_HOST
        opt = Options()
        self.lib.gpp_default_options(C.byref(opt))
        opt.precision = PRECISION.get(precision, precision)
        opt.semantics = SEMANTICS.get(semantics, semantics)
        opt.memory = memory
        opt.flags = flags
        opt.stream = stream
        opt.resonance_threshold = resonance_threshold
        opt.tile_ig = tile_ig
        opt.waves_per_block = waves_per_block
        return opt

    @staticmethod
    def _result(res: Result) -> dict:
        return {
            "achtemp": np.array([_cplx(z) for z in res.achtemp]),
            "asxtemp": np.array([_cplx(z) for z in res.asxtemp]),
            "achstemp": _cplx(res.achstemp),
            "unocc_kernel_ms": res.unocc_kernel_ms,
            "occ_kernel_ms": res.occ_kernel_ms,
            "total_ms": res.total_ms,
            "terms": res.terms,
            "fp64_fallback_terms": res.fp64_fallback_terms,
        }

    def sigma(self, inp: dict, precision="fp64", semantics="gppker",
              memory=None, flags=0, stream=None, resonance_threshold=0.0,
              tile_ig=0, waves_per_block=0, acht_n1_loc=False) -> dict:
        """One evaluation.  With acht_n1_loc=True the per-band log
        (gppKer.cpp:357) is returned too."""
        prob = make_problem(inp)
        opt = self._options(inp, precision, semantics, memory, flags, stream,
                            resonance_threshold, tile_ig, waves_per_block)
        res = Result()
        n1 = None
        if acht_n1_loc:
            n1 = np.zeros(max(prob.number_bands, 1), np.complex128)
            res.acht_n1_loc = n1.ctypes.data
        rc = self.lib.gpp_sigma_ex(self.h, C.byref(prob), C.byref(opt), C.byref(res))
        if rc != GPP_OK:
            raise GppError(rc, self.last_error())
        out = self._result(res)
        if n1 is not None:
            out["acht_n1_loc"] = n1[: prob.number_bands]
        return out

    def ff(self, inp: dict, semantics="cpp", memory=None, stream=None,
           parts=0) -> dict:
        """Full-frequency outputs (ff_sigma_ex): achsDtemp, the five
        [nfreqeval] arrays, and the kernel times and work counts."""
        opt = FFOptions()
        self.lib.ff_default_options(C.byref(opt))
        opt.semantics = FF_SEMANTICS[semantics]
        if memory is None:
            memory = MEM_DEVICE if hasattr(inp["aqsmtemp"], "data_ptr") else MEM_HOST
        opt.memory = memory
        opt.stream = stream
        opt.parts = parts
        ne = int(inp["nfreqeval"])
        outs = {k: np.zeros(ne, np.complex128) for k in FF_OUTPUTS}
        res = FFResult()
        for k in FF_OUTPUTS:
            setattr(res, k, outs[k].ctypes.data)
        prob = make_ff_problem(inp)
        rc = self.lib.ff_sigma_ex(self.h, C.byref(prob), C.byref(opt), C.byref(res))
        if rc != GPP_OK:
            raise GppError(rc, self.last_error())
        outs["achsDtemp"] = _cplx(res.achsDtemp)
        for k in ("slice_kernel_ms", "freq_kernel_ms", "total_ms", "slice_bytes",
                  "slices", "slice_terms", "freq_terms", "host_plan_ms",
                  "host_assemble_ms"):
            outs[k] = getattr(res, k)
        return outs

    def batch(self, inputs, precision="fp64", semantics="gppker",
              memory=None, flags=0, stream=None, resonance_threshold=0.0,
              tile_ig=0) -> list:
        """Many independent evaluations in one batched launch sequence
        (gpp_sigma_batch), replayed from a hipGraph while unchanged.
        `inputs` is a list of input dicts or a BatchProblems built from one
        (which skips rebuilding the problem table on every call)."""
        bp = inputs if isinstance(inputs, BatchProblems) else BatchProblems(inputs)
        n = len(bp)
        opt = self._options(bp.inputs[0], precision, semantics, memory, flags,
                            stream, resonance_threshold, tile_ig, 0)
        res = (Result * n)()
        rc = self.lib.gpp_sigma_batch(self.h, bp.table, n, C.byref(opt), res)
        if rc != GPP_OK:
            raise GppError(rc, self.last_error())
        return _batch_results(res, n)

    def autotune(self, inp: dict, precision="fp64", semantics="gppker",
                 memory=None, stream=None, resonance_threshold=0.0) -> tuple:
        """Run the tile autotuner for this problem's shape class; returns
        (tile_ig, waves_per_block), also cached in the handle."""
        prob = make_problem(inp)
        opt = self._options(inp, precision, semantics, memory, 0, stream,
                            resonance_threshold, 0, 0)
        k, w = C.c_int(), C.c_int()
        rc = self.lib.gpp_autotune(self.h, C.byref(prob), C.byref(opt),
                                   C.byref(k), C.byref(w))
        if rc != GPP_OK:
            raise GppError(rc, self.last_error())
        return k.value, w.value


FILL_VCOUL_LINEAR = 256


def fill_inputs(config: int, number_bands: int, ncouls: int,
                nodes_per_group: int, nvband: int, seed: int = 12345,
                lib: C.CDLL | None = None, vcoul_linear: bool = False) -> dict:
    """Synthetic inputs (numpy, host) from the library's generator
    (gpp_fill_inputs): config 1 constant fills, 2 canonical random,
    3 band-resolved energies with near-resonant wtilde; vcoul_linear
    replaces vcoul by vcoul[i] = i / ncouls."""
    lib = lib or load_library()
    flag = FILL_VCOUL_LINEAR if vcoul_linear else 0
    nb, nc = number_bands, ncouls
    ng = nc // nodes_per_group
    d = {
        "aqsmtemp": np.empty((nb, nc), np.complex128),
        "aqsntemp": np.empty((nb, nc), np.complex128),
        "I_eps_array": np.empty((ng, nc), np.complex128),
        "wtilde_array": np.empty((ng, nc), np.complex128),
        "vcoul": np.empty(nc, np.float64),
        "inv_igp_index": np.empty(max(ng, 1), np.int32),
        "indinv": np.empty(nc, np.int32),
        "wx_array": np.empty((nb, 3) if config == 3 else 3, np.float64),
    }
    rc = lib.gpp_fill_inputs(config | flag, seed, nb, nc, ng, nodes_per_group,
                             *[_ptr(d[k]) for k in (
                                 "aqsmtemp", "aqsntemp", "I_eps_array",
                                 "wtilde_array", "vcoul", "inv_igp_index",
                                 "indinv", "wx_array")])
    if rc != GPP_OK:
        raise GppError(rc, "gpp_fill_inputs failed")
    d["inv_igp_index"] = d["inv_igp_index"][:ng].copy()
    d.update(number_bands=nb, ncouls=nc, ngpown=ng, nvband=nvband)
    return d


def fill_inputs_device(config: int, number_bands: int, ncouls: int,
                       nodes_per_group: int, nvband: int, seed: int = 12345,
                       lib: C.CDLL | None = None,
                       vcoul_linear: bool = False) -> dict:
    """The same synthetic inputs generated in device memory (torch tensors on
    cuda:0, complex128 / float64 / int32), by gpp_fill_inputs_device."""
    import torch
    lib = lib or load_library()
    flag = FILL_VCOUL_LINEAR if vcoul_linear else 0
    nb, nc = number_bands, ncouls
    ng = nc // nodes_per_group
    dev = torch.device("cuda", 0)
    d = {
        "aqsmtemp": torch.empty((nb, nc), dtype=torch.complex128, device=dev),
        "aqsntemp": torch.empty((nb, nc), dtype=torch.complex128, device=dev),
        "I_eps_array": torch.empty((ng, nc), dtype=torch.complex128, device=dev),
        "wtilde_array": torch.empty((ng, nc), dtype=torch.complex128, device=dev),
        "vcoul": torch.empty(nc, dtype=torch.float64, device=dev),
        "inv_igp_index": torch.empty(max(ng, 1), dtype=torch.int32, device=dev),
        "indinv": torch.empty(nc, dtype=torch.int32, device=dev),
        "wx_array": torch.empty((nb, 3) if config == 3 else 3,
                                dtype=torch.float64, device=dev),
    }
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.gpp_fill_inputs_device(config | flag, seed, nb, nc, ng, nodes_per_group,
                                    *[d[k].data_ptr() for k in (
                                        "aqsmtemp", "aqsntemp", "I_eps_array",
                                        "wtilde_array", "vcoul",
                                        "inv_igp_index", "indinv", "wx_array")],
                                    stream)
    if rc != GPP_OK:
        raise GppError(rc, "gpp_fill_inputs_device failed")
    d["inv_igp_index"] = d["inv_igp_index"][:ng]
    d.update(number_bands=nb, ncouls=nc, ngpown=ng, nvband=nvband)
    return d


def _ff_arrays(nb, nc, ng, nf, alloc):
    return {
        "aqsmtemp": alloc((nb, nc), "c"), "aqsntemp": alloc((nb, nc), "c"),
        "I_epsR_array": alloc((nf, ng, nc), "c"),
        "I_epsA_array": alloc((nf, ng, nc), "c"),
        "vcoul": alloc((nc,), "d"), "inv_igp_index": alloc((max(ng, 1),), "i"),
        "indinv": alloc((nc,), "i"), "ekq": alloc((max(nb, 1),), "d"),
        "dFreqGrid": alloc((nf,), "d"), "pref": alloc((nf,), "d"),
        "dFreqBrd": alloc((nf,), "c"), "schDt_matrix": alloc((max(nb, 1), nf), "c"),
    }


def ff_fill_inputs(config: int, number_bands: int, nvband: int, ncouls: int,
                   ngpown: int, nFreq: int, nfreqeval: int, semantics="cpp",
                   seed: int = 2024, lib: C.CDLL | None = None) -> dict:
    """FF inputs (numpy, host) from the library's generator (ff_fill_inputs):
    config 1 the mini-app's fills, 2 seeded random."""
    lib = lib or load_library()
    dt = {"c": np.complex128, "d": np.float64, "i": np.int32}
    d = _ff_arrays(number_bands, ncouls, ngpown, nFreq,
                   lambda shape, t: np.empty(shape, dt[t]))
    scal = np.empty(4, np.float64)
    rc = lib.ff_fill_inputs(config, FF_SEMANTICS[semantics], seed, number_bands,
                            ncouls, ngpown, nFreq, *[_ptr(d[k]) for k in FF_ARRAYS],
                            _ptr(scal))
    if rc != GPP_OK:
        raise GppError(rc, "ff_fill_inputs failed")
    d["inv_igp_index"] = d["inv_igp_index"][:ngpown].copy()
    d.update(number_bands=number_bands, nvband=nvband, ncouls=ncouls,
             ngpown=ngpown, nFreq=nFreq, nfreqeval=nfreqeval,
             freqevalmin=scal[0], freqevalstep=scal[1], pref_zb=scal[2],
             occ=scal[3])
    return d


def ff_fill_inputs_device(config: int, number_bands: int, nvband: int,
                          ncouls: int, ngpown: int, nFreq: int, nfreqeval: int,
                          semantics="cpp", seed: int = 2024,
                          lib: C.CDLL | None = None) -> dict:
    """The same FF inputs generated in device memory (torch tensors on
    cuda:0) by ff_fill_inputs_device."""
    import torch
    lib = lib or load_library()
    dev = torch.device("cuda", 0)
    dt = {"c": torch.complex128, "d": torch.float64, "i": torch.int32}
    d = _ff_arrays(number_bands, ncouls, ngpown, nFreq,
                   lambda shape, t: torch.empty(shape, dtype=dt[t], device=dev))
    scal = np.empty(4, np.float64)
    rc = lib.ff_fill_inputs_device(config, FF_SEMANTICS[semantics], seed,
                                   number_bands, ncouls, ngpown, nFreq,
                                   *[d[k].data_ptr() for k in FF_ARRAYS],
                                   _ptr(scal), torch.cuda.current_stream().cuda_stream)
    if rc != GPP_OK:
        raise GppError(rc, "ff_fill_inputs_device failed")
    d["inv_igp_index"] = d["inv_igp_index"][:ngpown]
    d.update(number_bands=number_bands, nvband=nvband, ncouls=ncouls,
             ngpown=ngpown, nFreq=nFreq, nfreqeval=nfreqeval,
             freqevalmin=scal[0], freqevalstep=scal[1], pref_zb=scal[2],
             occ=scal[3])
    return d


# The batched workload of the benchmark: 64 independent band-resolved
# (config 3) instances, ncouls log-uniform in [2048, 32768], number_bands in
# {128, 256, 512, 1024}, nvband uniform in [2, 16], 20 G-vectors per G'
# column; and 512 small ones for the per-launch costs.
BATCH_WORKLOADS = {
    "batch": dict(n=64, nc=(2048, 32768), nb=(128, 256, 512, 1024), nv=(2, 16),
                  npg=20),
    "batch_small": dict(n=512, nc=(256, 2048), nb=(16, 32, 64), nv=(2, 8), npg=8),
}


def draw_batch(name: str, seed: int) -> list:
    """The instance shapes of a batch workload for `seed`:
    [(number_bands, nvband, ncouls, nodes_per_group, instance_seed), ...]."""
    w = BATCH_WORKLOADS[name]
    rng = np.random.default_rng(seed)
    out = []
    for i in range(w["n"]):
        nc = int(round(np.exp(rng.uniform(np.log(w["nc"][0]), np.log(w["nc"][1])))))
        nb = int(rng.choice(list(w["nb"])))
        nv = int(rng.integers(w["nv"][0], w["nv"][1] + 1))
        out.append((nb, nv, nc, w["npg"], seed + 1000 + i))
    return out


def column_slice(inp: dict, ng: int) -> dict:
    """The sub-problem of the first `ng` G' columns (my_igp < ng): the full
    sums are sums over columns, so a column slice is checked on its own."""
    d = dict(inp)
    d["I_eps_array"] = inp["I_eps_array"][:ng]
    d["wtilde_array"] = inp["wtilde_array"][:ng]
    d["inv_igp_index"] = inp["inv_igp_index"][:ng]
    d["ngpown"] = ng
    return d


def band_slice(inp: dict, n1_begin: int, n1_end: int) -> dict:
    """The sub-problem of bands [n1_begin, n1_end) (band additivity: the full
    sums are the sums of the sub-problems')."""
    d = dict(inp)
    d["aqsmtemp"] = inp["aqsmtemp"][n1_begin:n1_end]
    d["aqsntemp"] = inp["aqsntemp"][n1_begin:n1_end]
    if inp["wx_array"].ndim == 2:
        d["wx_array"] = inp["wx_array"][n1_begin:n1_end]
    d["number_bands"] = n1_end - n1_begin
    d["nvband"] = max(0, min(inp["nvband"], n1_end) - n1_begin)
    return d
