"""Writes the golden fixtures of tests/golden/ from the reference's own code.

Run in a container that has the reference tree at /root/reference:

    python tests/golden/make_golden.py

1. ref_constfill.json — the achtemp lines printed by the reference mini-app
   gppKerOpenMPFor_xl.cpp (gppKer.cpp without the Intel ITT calls and with
   the non-occupied mask of :335-338 active, which keeps every term of the
   constant fills), built
   unmodified by oracle/Makefile, for several shapes with its own constant
   fills (gppKer.cpp:241-264).  OMP_NUM_THREADS=1 so that the printed sums
   are in one fixed order.
2. ref_occ_kat_<variant>.json — seeded random inputs and, for every
   (n1, my_igp, iw), the ssxt/scht sums returned by the reference's own
   flagOCC_solver (BGW-Kernels/gppKer.cpp:130 and gppKerOpenMPFor_xl.cpp:125,
   the two I_eps index variants), through oracle/ref_kat.cpp.
   Two input sets: "branches" draws wtilde so that every branch of
   ssxt_scht_solver (gppKer.cpp:35-55) is taken; "masktrue" keeps
   |wx - wtilde|^2 > 0.25 everywhere, where the occupied solver's scht equals
   the non-occupied sum of gppKer.cpp:326-343 term by term.
3. ref_ff_constfill.json — the result lines printed by the full-frequency
   mini-app FFKer_cori/ffKernel.cpp, built unmodified, for several shapes
   with its own fills (ffKernel.cpp:385-432).
4. ref_ff_kat.json — seeded random inputs and the outputs of the
   reference's own achsDtemp_Kernel, asxDtemp_Kernel, achDtemp_Kernel and
   achDtemp_cor_Kernel (ffKernel.cpp:107,133,155,216) through
   oracle/ff_ref_kat.cpp, OMP_NUM_THREADS=1.
The fixtures hold data only: inputs and the reference's outputs.
"""
import json
import os
import re
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
OUT = ROOT / "tests" / "golden"
ORACLE = ROOT / "oracle"

SHAPES = [  # number_bands nvband ncouls nodes_per_group
    (64, 2, 2048, 8),
    (8, 2, 64, 4),
    (16, 3, 200, 16),
    (5, 5, 40, 8),
    (12, 0, 100, 7),
]


def build():
    subprocess.run(["make", "-C", str(ORACLE), "ref"], check=True)


def parse_c(s):
    m = re.match(r"\(([^,]+),([^)]+)\)", s.strip())
    return [float(m.group(1)), float(m.group(2))]


def constfill():
    recs = []
    env = dict(os.environ, OMP_NUM_THREADS="1")
    for nb, nv, nc, npg in SHAPES:
        out = subprocess.run([str(ORACLE / "_ref" / "gppKer_xl"), str(nb),
                              str(nv), str(nc), str(npg)], env=env,
                             check=True, capture_output=True, text=True).stdout
        ach = {}
        for line in out.splitlines():
            m = re.match(r"achtemp\[(\d)\] = (.*)", line)
            if m:
                ach[int(m.group(1))] = parse_c(m.group(2))
        recs.append({"number_bands": nb, "nvband": nv, "ncouls": nc,
                     "nodes_per_group": npg,
                     "achtemp": [ach[0], ach[1], ach[2]],
                     "printed": [l for l in out.splitlines()
                                 if l.startswith("achtemp")]})
    (OUT / "ref_constfill.json").write_text(json.dumps(recs, indent=1))


def kat_inputs(kind, seed):
    rng = np.random.default_rng(seed)
    nb, nc, ng = 3, 40, 6
    cx = lambda *s: (rng.normal(size=s) + 1j * rng.normal(size=s)) / np.sqrt(nc)
    aqsm, aqsn = cx(nb, nc), cx(nb, nc)
    ieps = rng.uniform(0, 1, (ng, nc)) + 1j * rng.uniform(-0.1, 0.1, (ng, nc))
    wx = np.array([[3.0, 4.0, 5.0], [2.2, 3.1, -1.5], [0.7, 1.0, 1.3]])
    if kind == "masktrue":
        wx = np.array([[3.0, 4.0, 5.0], [3.5, 4.5, 5.5], [2.9, 6.0, 8.0]])
        wt = rng.uniform(0.5, 2.0, (ng, nc)) + 1j * rng.uniform(-0.05, 0.05, (ng, nc))
    else:
        wt = rng.uniform(0.5, 3.0, (ng, nc)) + 1j * rng.uniform(-0.05, 0.05, (ng, nc))
        sel = rng.uniform(size=(ng, nc))
        # near a frequency: |wdiff|^2 <= 0.25 (second branch of the solver)
        pick = wx.ravel()[rng.integers(0, 9, (ng, nc))]
        near = sel < 0.25
        wt[near] = pick[near] + rng.uniform(-0.3, 0.3, near.sum()) \
            + 1j * rng.uniform(-0.05, 0.05, near.sum())
        # tiny wtilde: |delw|^2 <= to1 (third branch)
        tiny = (sel >= 0.25) & (sel < 0.35)
        wt[tiny] = 1e-5 * (rng.uniform(-1, 1, tiny.sum())
                           + 1j * rng.uniform(-1, 1, tiny.sum()))
    igp = np.array([0, 7, 14, 21, 39, 33], np.int32)
    return dict(nb=nb, nc=nc, ng=ng, aqsm=aqsm, aqsn=aqsn, ieps=ieps, wt=wt,
                wx=wx, igp=igp)


def run_kat(exe, d):
    blob = (np.array([d["nb"], d["nc"], d["ng"]], np.int32).tobytes()
            + d["aqsm"].astype(np.complex128).tobytes()
            + d["aqsn"].astype(np.complex128).tobytes()
            + d["ieps"].astype(np.complex128).tobytes()
            + d["wt"].astype(np.complex128).tobytes()
            + d["wx"].astype(np.float64).tobytes()
            + d["igp"].astype(np.int32).tobytes())
    out = subprocess.run([str(exe)], input=blob, capture_output=True,
                         check=True).stdout.decode()
    rows = []
    for line in out.splitlines():
        f = line.split()
        rows.append([int(f[0]), int(f[1]), int(f[2])] + [float(x) for x in f[3:]])
    return rows


def cl(a):
    a = np.asarray(a)
    return np.stack([a.real, a.imag], -1).tolist()


def kat():
    for kind, seed in (("branches", 7), ("masktrue", 11)):
        d = kat_inputs(kind, seed)
        rec = {"kind": kind, "seed": seed, "number_bands": d["nb"],
               "ncouls": d["nc"], "ngpown": d["ng"],
               "aqsmtemp": cl(d["aqsm"]), "aqsntemp": cl(d["aqsn"]),
               "I_eps_array": cl(d["ieps"]), "wtilde_array": cl(d["wt"]),
               "wx_array": d["wx"].tolist(), "igp": d["igp"].tolist()}
        for variant in ("bgwk", "xl"):
            rec["rows_" + variant] = run_kat(ORACLE / "_ref" / f"occ_kat_{variant}", d)
        (OUT / f"ref_occ_kat_{kind}.json").write_text(json.dumps(rec))


FF_SHAPES = [  # number_bands nvband ncouls ngpown nFreq nfreqeval
    (4, 2, 64, 8, 6, 5),
    (14, 13, 40, 5, 9, 3),
    (3, 0, 30, 3, 4, 2),
    (12, 3, 50, 6, 4, 4),
]


def ff_constfill():
    recs = []
    env = dict(os.environ, OMP_NUM_THREADS="1")
    for shape in FF_SHAPES:
        out = subprocess.run([str(ORACLE / "_ref" / "ffKer_cpp")] +
                             [str(x) for x in shape], env=env, check=True,
                             capture_output=True, text=True).stdout
        vals = {}
        for line in out.splitlines():
            m = re.match(r"(achsDtemp|asxDtemp|achDtemp_cor) = (.*)", line)
            if m:
                vals[m.group(1)] = parse_c(m.group(2))
        recs.append(dict(zip(("number_bands", "nvband", "ncouls", "ngpown",
                              "nFreq", "nfreqeval"), shape), **vals))
    (OUT / "ref_ff_constfill.json").write_text(json.dumps(recs, indent=1))


def ff_kat_inputs(seed, nb=6, nv=3, nc=24, ng=4, nf=7, ne=6):
    """Random FF inputs: a non-uniform increasing grid, band energies that
    put wx below zero, inside and beyond the grid, strided G' map."""
    rng = np.random.default_rng(seed)
    u = lambda *s: rng.random(s)
    cx = lambda *s: (2 * u(*s) - 1) + 1j * (2 * u(*s) - 1)
    grid = np.concatenate([[0.0], np.cumsum(1.5 + u(nf - 1))])
    pref = np.empty(nf)
    pref[:-1] = np.diff(grid) / 3.14
    pref[-1] = pref[-2]
    pref[0] *= 0.5
    pref[-1] *= 0.5
    ekq = np.sort(-14.0 + 24.0 * u(nb))
    return dict(nb=nb, nv=nv, nc=nc, ng=ng, nf=nf, ne=ne,
                fmin=0.25, fstep=0.5, pref_zb=0.5 / 3.14, occ=1.0,
                aqsm=cx(nb, nc) / np.sqrt(nc), aqsn=cx(nb, nc) / np.sqrt(nc),
                er=cx(nf, ng, nc), ea=cx(nf, ng, nc),
                vcoul=0.5 + u(nc), inv=np.arange(ng, dtype=np.int32) * (nc // ng),
                indinv=rng.permutation(nc).astype(np.int32), ekq=ekq,
                grid=grid, pref=pref, brd=0.5 * u(nf) + 1j * (0.05 + 0.1 * u(nf)),
                schdt=cx(nb, nf))


def ff_run_kat(d):
    blob = (np.array([d["nb"], d["nv"], d["nc"], d["ng"], d["nf"], d["ne"]],
                     np.int32).tobytes()
            + np.array([d["fmin"], d["fstep"], d["pref_zb"], d["occ"]]).tobytes()
            + b"".join(np.ascontiguousarray(d[k], np.complex128).tobytes()
                       for k in ("aqsm", "aqsn", "er", "ea"))
            + d["vcoul"].astype(np.float64).tobytes()
            + d["inv"].astype(np.int32).tobytes()
            + d["indinv"].astype(np.int32).tobytes()
            + b"".join(d[k].astype(np.float64).tobytes()
                       for k in ("ekq", "grid", "pref"))
            + b"".join(np.ascontiguousarray(d[k], np.complex128).tobytes()
                       for k in ("brd", "schdt")))
    env = dict(os.environ, OMP_NUM_THREADS="1")
    out = subprocess.run([str(ORACLE / "_ref" / "ff_kat")], input=blob, env=env,
                         capture_output=True, check=True).stdout.decode()
    lines = out.splitlines()
    achs = [float(x) for x in lines[0].split()[1:]]
    rows = [[float(x) for x in l.split()[1:]] for l in lines[1:]]
    return achs, rows


def ff_kat():
    recs = []
    for seed in (5, 6, 7):
        d = ff_kat_inputs(seed)
        achs, rows = ff_run_kat(d)
        rec = {"seed": seed}
        for k, v in d.items():
            if isinstance(v, np.ndarray):
                rec[k] = cl(v) if np.iscomplexobj(v) else v.tolist()
            else:
                rec[k] = v
        rec["achsDtemp"] = achs
        rec["asxDtemp"] = [r[0:2] for r in rows]
        rec["achDtemp"] = [r[2:4] for r in rows]
        rec["achDtemp_cor"] = [r[4:6] for r in rows]
        recs.append(rec)
    (OUT / "ref_ff_kat.json").write_text(json.dumps(recs))


if __name__ == "__main__":
    if not Path("/root/reference").exists():
        sys.exit("needs the reference tree at /root/reference")
    build()
    constfill()
    kat()
    ff_constfill()
    ff_kat()
    print("wrote", sorted(p.name for p in OUT.glob("*.json")))
