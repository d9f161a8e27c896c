"""Shared test fixtures.

The product package lives in the directory `fortkernels-_amd/` (not a valid
module name), so it is loaded here by path under the name `gppk`.  The CPU
oracle (oracle/oracle.py) is loaded as `gpp_oracle`; it is the checker only.
"""
import importlib.util
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"


def _load(name, path, pkg=False):
    if name in sys.modules:
        return sys.modules[name]
    kw = {"submodule_search_locations": [str(path.parent)]} if pkg else {}
    spec = importlib.util.spec_from_file_location(name, path, **kw)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


gppk = _load("gppk", ROOT / "fortkernels-_amd" / "__init__.py", pkg=True)
gpp_oracle = _load("gpp_oracle", ROOT / "oracle" / "oracle.py")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def handle():
    if not gpu_available():
        pytest.skip("no GPU")
    h = gppk.Handle(0)
    yield h
    h.close()


def rel_err(a, b):
    """max_i |a_i - b_i| / |b_i| over entries with b_i != 0 (else absolute)."""
    import numpy as np
    a = np.atleast_1d(np.asarray(a, np.complex128))
    b = np.atleast_1d(np.asarray(b, np.complex128))
    d = np.abs(a - b)
    s = np.abs(b)
    return float(np.max(np.where(s > 0, d / np.where(s > 0, s, 1), d)))
