"""The two largest workloads at their full sizes on the GPU, checked against
the oracle on slices it finishes in seconds and through band additivity:

  * the large instance 2048/16/65536/20 (band-resolved energies, 5%
    near-resonant wtilde, 11.2 GB generated in device memory), FP64 and
    mixed precision;
  * the batch of 64 band-resolved instances (ncouls log-uniform in
    [2048, 32768], number_bands in {128, 256, 512, 1024}, nvband in [2, 16],
    25.3 GB), autotuned tile, FP64 and mixed.

Both scale the loop of gppKer.cpp:270-361."""
import numpy as np
import pytest

from conftest import gppk, gpp_oracle as orc, rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-11


def say(*a):
    print(*a, flush=True)


def to_host(d: dict) -> dict:
    return {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in d.items()}


def slice_check(handle, dev: dict, nb: int, ng: int):
    """The first `nb` bands x `ng` G' columns on the GPU against the oracle
    on a host copy of the same slice, FP64."""
    sub = gppk.column_slice(gppk.band_slice(dev, 0, nb), min(ng, dev["ngpown"]))
    got = handle.sigma(sub)
    ref = orc.sigma(to_host(sub), want_achstemp=False)
    for k in ("achtemp", "asxtemp"):
        e = rel_err(got[k], ref[k])
        assert e <= FP64_TOL, (k, e, got[k], ref[k])
    return got


def entry_err(a: dict, b: dict) -> float:
    return max(rel_err(a[k], b[k]) for k in ("achtemp", "asxtemp"))


# Mixed-precision bound per achtemp/asxtemp entry against FP64 on these
# workloads (measured 1.6e-8 on the large instance and 4.2e-7 on the batch
# of these seeds; the bounds are 2-3x that).
MIXED_LARGE_BOUND = 5e-8
MIXED_BATCH_BOUND = 1e-6


@pytest.mark.parametrize("cfg,bound", [(2, 5e-7), (3, 4e-8)])
def test_canonical_shape_mixed_full(handle, cfg, bound):
    """512/2/32768/20 in full, mixed against FP64 per entry (measured
    2.3e-7 with the canonical distributions, 1.9e-8 band-resolved); on the
    band-resolved inputs a resonance window of 4e-3 brings every entry below
    1e-8 (measured 5.5e-9)."""
    import torch
    dev = gppk.fill_inputs_device(cfg, 512, 32768, 20, nvband=2, seed=2024)
    full = handle.sigma(dev)
    mixed = handle.sigma(dev, precision="mixed")
    e = entry_err(mixed, full)
    say(f"canonical shape config {cfg}: mixed vs fp64 per entry {e:.3e}")
    assert e <= bound
    if cfg == 3:
        win = handle.sigma(dev, precision="mixed", resonance_threshold=4e-3)
        e = entry_err(win, full)
        say(f"canonical shape config 3, window 4e-3: {e:.3e}")
        assert e <= 1e-8
    del dev
    torch.cuda.empty_cache()


def test_large_instance_2048_16_65536_20(handle):
    import torch
    say("large: generating 2048/16/65536/20 on the device")
    dev = gppk.fill_inputs_device(3, 2048, 65536, 20, nvband=16, seed=77)
    torch.cuda.synchronize()
    say("large: oracle slice 17 bands x 128 columns")
    slice_check(handle, dev, 17, 128)
    say("large: full FP64 and band additivity")
    full = handle.sigma(dev)
    parts = [handle.sigma(gppk.band_slice(dev, a, b))
             for a, b in ((0, 700), (700, 2048))]
    for k in ("achtemp", "asxtemp"):
        assert rel_err(parts[0][k] + parts[1][k], full[k]) <= 1e-12, k
    say("large: mixed")
    mixed = handle.sigma(dev, precision="mixed")
    e = entry_err(mixed, full)
    say(f"large: mixed vs fp64 per entry {e:.3e}")
    assert e <= MIXED_LARGE_BOUND
    # occupied bands run in FP64 in both paths
    assert rel_err(mixed["asxtemp"], full["asxtemp"]) <= 1e-13
    del dev, full, parts, mixed
    torch.cuda.empty_cache()


def test_batch_of_64_instances(handle):
    import torch
    shapes = gppk.draw_batch("batch", 2024)
    say(f"batch: generating {len(shapes)} instances on the device")
    insts = [gppk.fill_inputs_device(3, nb, nc, npg, nvband=nv, seed=s)
             for nb, nv, nc, npg, s in shapes]
    torch.cuda.synchronize()
    big = max(insts, key=lambda d: d["number_bands"] * d["ngpown"] * d["ncouls"])
    k, w = handle.autotune(big)
    say(f"batch: autotuned tile K={k} W={w}")
    table = gppk.BatchProblems(insts)
    res = handle.batch(table)                     # the tuned tile
    say("batch: one by one at the same tile")
    for i, d in enumerate(insts):
        one = handle.sigma(d, tile_ig=k, waves_per_block=w)
        for key in ("achtemp", "asxtemp"):
            assert np.array_equal(res[i][key], one[key]), (i, key)
        assert res[i]["achstemp"] == one["achstemp"], i
    say("batch: oracle slices (nvband+1 bands x 128 columns of each)")
    for (nb, nv, nc, npg, s), d in zip(shapes, insts):
        slice_check(handle, d, nv + 1, 128)
    say("batch: mixed")
    mixed = handle.batch(table, precision="mixed")
    e = max(entry_err(m, r) for m, r in zip(mixed, res))
    say(f"batch: mixed vs fp64 per entry {e:.3e}")
    assert e <= MIXED_BATCH_BOUND
    # a replay of the captured graph gives the same bits
    again = handle.batch(table)
    for a, b in zip(again, res):
        assert np.array_equal(a["achtemp"], b["achtemp"])
    del insts, table
    torch.cuda.empty_cache()
