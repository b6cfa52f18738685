"""The config-driven driver (ns_spectral_run) on the GPU: a Taylor-Green run
from an INI file whose .dat / VTK / PVD / checkpoint outputs are parsed back
and checked against the closed-form decay omega = 2 cos x cos y exp(-2 nu t),
E = 0.25 exp(-4 nu t); a restart from its checkpoint; and a short forced FP32
turbulence run from keys given on the command line."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu] {request.node.name}", flush=True)


def run(ns, args, cwd):
    exe = ns.DRIVER_PATH
    assert os.path.exists(exe), "ns_spectral_run not built"
    r = subprocess.run([exe] + args, capture_output=True, text=True, cwd=cwd, timeout=120)
    print(r.stdout[-1500:], r.stderr[-800:], flush=True)
    return r


TG_INI = """[solver]
type = spectral
[grid]
nx = 64
ny = 64
[fluid]
Re = 100.0
[time]
dt = 0.001
t_end = 0.1
adaptive = 0
[init]
type = taylor_green
[output]
interval = 50
format = dat,vtk,vtu
diagnostics_interval = 25
[checkpoint]
interval = 50
file = tg.h5
"""


def test_taylor_green_run_outputs_and_restart(ns, tmp_path):
    (tmp_path / "tg.ini").write_text(TG_INI)
    r = run(ns, ["tg.ini"], tmp_path)
    assert r.returncode == 0
    assert "Total steps: 100" in r.stdout
    E = float(re.search(r"Final energy: (\S+)", r.stdout).group(1))
    assert E == pytest.approx(0.25 * np.exp(-4 * 0.01 * 0.1), rel=1e-6)
    for f in ("spectral_solution_step_0050.dat", "spectral_solution_final.dat",
              "spectral_solution_000050.vtk", "spectral_solution_000100.vtr",
              "spectral_solution.pvd", "tg.h5"):
        assert (tmp_path / f).exists(), f
    dat = np.loadtxt(tmp_path / "spectral_solution_final.dat")
    assert dat.shape == (64 * 64, 6)
    x, y, w = dat[:, 0], dat[:, 1], dat[:, 4]
    exact = 2 * np.cos(x) * np.cos(y) * np.exp(-2 * 0.01 * 0.1)
    assert np.abs(w - exact).max() < 1e-7
    # restart from the checkpoint written at t = 0.1 and run to t = 0.15
    r = run(ns, ["tg.ini", "checkpoint.restart=tg.h5", "time.t_end=0.15", "checkpoint.interval=0",
                 "output.prefix=restarted", "output.format=dat"], tmp_path)
    assert r.returncode == 0 and "Restarted from tg.h5" in r.stdout
    assert "Total steps: 50" in r.stdout
    dat = np.loadtxt(tmp_path / "restarted_final.dat")
    exact = 2 * np.cos(dat[:, 0]) * np.cos(dat[:, 1]) * np.exp(-2 * 0.01 * 0.15)
    assert np.abs(dat[:, 4] - exact).max() < 1e-7


def test_forced_fp32_run_from_keys(ns, tmp_path):
    args = ["grid.nx=256", "grid.ny=256", "fluid.Re=10000", "fluid.drag=0.1",
            "spectral.precision=fp32", "spectral.scheme=if_rk4",
            "spectral.hyperviscosity.coefficient=" + repr(float(1e-16 * (np.sqrt(2) * 128) ** 8)),
            "spectral.hyperviscosity.order=8", "spectral.hyperviscosity.cutoff_ratio=0",
            "forcing.type=white", "forcing.kmin=4", "forcing.kmax=6", "forcing.seed=42",
            "forcing.injection_rate=1e-3", "init.type=random_phase", "init.seed=42",
            "time.cfl_factor=0.5", "time.t_end=100", "time.max_steps=300",
            "output.interval=0", "output.format=none", "output.diagnostics_interval=100",
            "output.resolution_check=0"]
    r = run(ns, args, tmp_path)
    assert r.returncode == 0
    assert "Total steps: 300" in r.stdout
    E = float(re.search(r"Final energy: (\S+)", r.stdout).group(1))
    assert 0.5 < E < 1.0
