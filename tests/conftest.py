"""Shared test helpers: loaders for the package (directory gp-mood_amd/), the
CPU oracle (oracle/, test infrastructure) and the golden fixtures."""
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


def pytest_collection_modifyitems(session, config, items):
    """Run the Fortran drop-in cases first.  They start the drop-in program
    as a child process.  The GPU grants the resource of cooperative launches
    to one process at a time, and a test process that has made one keeps it
    until it exits; the child's MOOD kernel would then fall back to an
    ordinary launch (same results, covered by
    test_run_persistent_launch_forms).  Running the drop-ins first keeps them
    on the cooperative path that a drop-in run alone takes."""
    items.sort(key=lambda it: 0 if "test_gpu_dropin" in it.nodeid else 1)


def load_package():
    if "gpmood_amd" in sys.modules:
        return sys.modules["gpmood_amd"]
    pkg_dir = os.path.join(ROOT, "gp-mood_amd")
    spec = importlib.util.spec_from_file_location("gpmood_amd", os.path.join(pkg_dir, "__init__.py"),
                                                  submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gpmood_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def load_oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        import oracle as O
    finally:
        sys.path.pop(0)
    return O


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def tables_from_npz(z):
    return {k: z[k] for k in ("zT", "d2x", "d2y", "gw", "ixiy")}


def case_options(z):
    """Grid, domain and step options of a golden step case (defaults: square
    unit domain, SSP_RK3, HLLC, no source term)."""
    n = int(z["n"])
    return dict(nf=int(z["nf"]) if "nf" in z else n, lx=float(z["lx"]) if "lx" in z else 1.0,
                ly=float(z["ly"]) if "ly" in z else 1.0, flux=str(z["flux"]) if "flux" in z else "hllc",
                integrator=str(z["integrator"]) if "integrator" in z else "ssp_rk3",
                problem=int(z["problem"]) if "problem" in z else 0,
                method=str(z["method"]) if "method" in z else "gp_mood",
                cascade=str(z["cascade"]) if "cascade" in z else "reference")


@pytest.fixture(scope="session")
def gpm():
    return load_package()


@pytest.fixture(scope="session")
def orc():
    O = load_oracle()
    O.lib()  # raises if not built
    return O


def has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
