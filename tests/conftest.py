"""Shared fixtures: the product package (navier-stokes-solvers_amd, loaded by
path because its directory name is not an identifier) and the CPU oracle."""
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "navier-stokes-solvers_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_package():
    name = "ns_spectral_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def ns():
    return load_package()


@pytest.fixture(scope="session")
def oc():
    import oracle
    oracle.lib()
    return oracle


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0)


def rk4_one_step_errors(step, w0, dts, substeps=64):
    """One-step error of `step(w, dt, m)` (m steps of dt/m from w) against the
    same system stepped with `substeps` steps, for each dt."""
    errs = []
    for dt in dts:
        one, fine = step(w0, dt, 1), step(w0, dt, substeps)
        errs.append(np.linalg.norm(one - fine) / np.linalg.norm(fine))
    return np.array(errs)


def quadratic_invariants(w, k2):
    """Parseval sums proportional to E and Z of a half spectrum (kx = 0 once)."""
    wt = np.ones_like(k2)
    wt[:, 1:] = 2.0
    a2 = np.abs(w) ** 2
    return (np.sum(wt * a2 * np.where(k2 > 0, 1.0 / np.maximum(k2, 1e-300), 0.0)),
            np.sum(wt * a2))

