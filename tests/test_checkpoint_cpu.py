"""Checkpoint file format on the CPU: host arrays of an NSSpectralData written
and read back through the library's HDF5 writer/reader (no device needed: the
device half of the struct is absent), and the file layout of the reference's
src/common/checkpoint.c:266-302 checked with h5dump where it is installed."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest


def host_struct(ns, nx, ny, rng):
    d = ns.NSSpectralData()
    d.nx, d.ny, d.nkx, d.nky = nx, ny, nx // 2 + 1, ny
    d.Lx, d.Ly, d.Re, d.dt, d.t, d.cfl_factor = 2 * np.pi, 3.0, 1234.5, 1e-3, 0.75, 0.4
    keep = []
    for name in ("u", "v", "p", "omega", "psi"):
        a = rng.standard_normal((ny, nx))
        keep.append(a)
        setattr(d, name, a.ctypes.data_as(ns.c_double_p))
    for name in ("u_hat", "v_hat", "omega_hat", "psi_hat", "sfd_hat"):
        a = rng.standard_normal((ny, (nx // 2 + 1) * 2))
        keep.append(a)
        setattr(d, name, a.ctypes.data_as(ns.c_double_p))
    d.device = None
    return d, keep


def hdf5_available():
    return any(os.path.exists(p) for p in ("/opt/conda/lib/libhdf5.so",)) or \
        shutil.which("h5dump") is not None


@pytest.mark.skipif(not hdf5_available(), reason="no HDF5 library on this host")
def test_checkpoint_round_trip_and_layout(ns, tmp_path):
    L = ns.load()
    rng = np.random.default_rng(5)
    a, keep_a = host_struct(ns, 32, 16, rng)
    fn = str(tmp_path / "ck.h5").encode()
    assert L.ns_checkpoint_write_spectral(fn, C.byref(a)) == 0, L.ns_checkpoint_error()
    b, keep_b = host_struct(ns, 32, 16, np.random.default_rng(6))
    b.t = b.dt = b.Re = 0.0
    assert L.ns_checkpoint_read_spectral(fn, C.byref(b)) == 0, L.ns_checkpoint_error()
    assert (b.t, b.dt, b.Re, b.Lx, b.Ly, b.cfl_factor) == (a.t, a.dt, a.Re, a.Lx, a.Ly, a.cfl_factor)
    for i, name in enumerate(("u", "v", "omega", "psi")):
        j = ("u", "v", "p", "omega", "psi").index(name)
        assert np.array_equal(keep_a[j], keep_b[j]), name
    for k in range(4):  # u_hat, v_hat, omega_hat, psi_hat (sfd off)
        assert np.array_equal(keep_a[5 + k], keep_b[5 + k])
    # a grid mismatch is refused
    c, keep_c = host_struct(ns, 16, 16, rng)
    assert L.ns_checkpoint_read_spectral(fn, C.byref(c)) != 0
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    if os.path.exists(h5dump):
        out = subprocess.run([h5dump, "-H", fn.decode()], capture_output=True, text=True).stdout
        attrs = [l.split('"')[1] for l in out.splitlines() if l.strip().startswith("ATTRIBUTE")]
        # h5dump lists attributes by name; the writer creates them in the
        # reference's order (solver, time, dt, Re, nx, ny, Lx, Ly)
        assert {"solver", "time", "dt", "Re", "nx", "ny", "Lx", "Ly"} <= set(attrs)
        assert "forcing_calls" in attrs and 'GROUP "fields"' in out
        for ds, shape in (("u", "( 16, 32 )"), ("omega_hat", "( 16, 17, 2 )")):
            i = out.index(f'DATASET "{ds}"')
            assert shape in out[i:i + 300], ds
