"""Output formats and the config-driven driver on the CPU.

The VTK writers (the formats of the reference's src/output/vtk_writer.c:10-135
and the PVD index of :273-330) are exercised on host arrays of an
NSSpectralData (no device needed) and parsed back; the driver's refusals
(finite-difference solver, no device) are checked without a GPU.  The
reference's writer cannot be compiled here (its header pulls in fftw3.h), so
the format is pinned by parsing: structure lines, counts and values."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest


def host_struct(ns, nx, ny, rng):
    d = ns.NSSpectralData()
    d.nx, d.ny, d.nkx, d.nky = nx, ny, nx // 2 + 1, ny
    keep = {}
    keep["x"] = 2 * np.pi * np.arange(nx) / nx
    keep["y"] = 2 * np.pi * np.arange(ny) / ny
    for name in ("u", "v", "omega"):
        keep[name] = rng.standard_normal((ny, nx))
    for k, a in keep.items():
        setattr(d, k, a.ctypes.data_as(ns.c_double_p))
    return d, keep


def test_vtk_legacy_format(ns, tmp_path):
    L = ns.load()
    d, k = host_struct(ns, 8, 4, np.random.default_rng(1))
    fn = str(tmp_path / "a.vtk")
    assert L.ns_vtk_write_spectral_legacy(fn.encode(), C.byref(d), 0.125) == 0
    lines = open(fn).read().splitlines()
    assert lines[:6] == ["# vtk DataFile Version 3.0", "Navier-Stokes Solution t=0.125000",
                         "ASCII", "DATASET STRUCTURED_GRID", "DIMENSIONS 8 4 1",
                         "POINTS 32 double"]
    pts = np.array([l.split() for l in lines[6:38]], float)
    assert np.allclose(pts[:, 0], np.tile(k["x"], 4), atol=5e-10)
    assert np.allclose(pts[:, 1], np.repeat(k["y"], 8), atol=5e-10)
    i = lines.index("VECTORS velocity double")
    vec = np.array([l.split() for l in lines[i + 1:i + 33]], float)
    assert np.allclose(vec[:, 0], k["u"].ravel(), atol=5e-10)
    assert np.allclose(vec[:, 1], k["v"].ravel(), atol=5e-10)
    i = lines.index("SCALARS vorticity double 1")
    assert lines[i + 1] == "LOOKUP_TABLE default"
    w = np.array(lines[i + 2:i + 34], float)
    assert np.allclose(w, k["omega"].ravel(), atol=5e-10)
    assert "TIME 1 1 double" in lines and "POINT_DATA 32" in lines


def test_vtk_xml_and_pvd(ns, tmp_path):
    import xml.etree.ElementTree as ET
    L = ns.load()
    d, k = host_struct(ns, 8, 4, np.random.default_rng(2))
    pvd = str(tmp_path / "run.pvd")
    for step, t in enumerate((0.0, 0.5)):
        fn = str(tmp_path / f"s{step}.vtr")
        assert L.ns_vtk_write_spectral_xml(fn.encode(), C.byref(d), t) == 0
        assert L.ns_vtk_update_pvd(pvd.encode(), f"s{step}.vtr".encode(), t) == 0
    root = ET.parse(fn).getroot()
    grid = root.find("RectilinearGrid")
    assert root.get("type") == "RectilinearGrid" and grid.get("WholeExtent") == "0 7 0 3 0 0"
    arrays = {a.get("Name"): np.array(a.text.split(), float)
              for a in root.iter("DataArray")}
    assert arrays["TimeValue"][0] == 0.5
    assert np.allclose(arrays["vorticity"], k["omega"].ravel(), atol=5e-10)
    vel = arrays["velocity"].reshape(-1, 3)
    assert np.allclose(vel[:, 0], k["u"].ravel(), atol=5e-10) and np.all(vel[:, 2] == 0)
    assert np.allclose(arrays["X"], k["x"], atol=5e-10) and np.allclose(arrays["Y"], k["y"], atol=5e-10)
    coll = ET.parse(pvd).getroot().find("Collection")
    ds = [(e.get("timestep"), e.get("file")) for e in coll.findall("DataSet")]
    assert ds == [("0.000000000", "s0.vtr"), ("0.500000000", "s1.vtr")]


def test_driver_refusals_without_device(ns, tmp_path):
    exe = ns.DRIVER_PATH
    if not os.path.exists(exe):
        pytest.skip("driver not built")
    r = subprocess.run([exe, "solver.type=finite_difference"], capture_output=True, text=True,
                       cwd=tmp_path, timeout=60)
    assert r.returncode == 2 and "only the spectral solver" in r.stderr
    r = subprocess.run([exe, "no_such_file.ini"], capture_output=True, text=True, cwd=tmp_path,
                       timeout=60)
    assert r.returncode == 2
