"""The allocation-free forward transform, nsd_physical_to_spectral_noalloc,
against the context path's nsd_physical_to_spectral (the same r2c, 1/(nx ny)
scaling and 2/3 mask of spectral_transforms.c:9-22), through the C-ABI.

Device buffers come from the HIP runtime the library itself links.  Two
calls in a row on one GPU share its scratch array: the second, smaller grid
must not see anything of the first.  FP64 tolerance 1e-14 relative L2 (the
two paths run the same kernels; only the twiddle tables are computed on the
device here and on the host there, one rounding apart); masked modes are
exactly zero."""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
H2D, D2H = 1, 2


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu] {request.node.name}", flush=True)


def hip_runtime(ns):
    ns.load()  # the runtime libnsspectral.so links is loaded first; this is the same one
    H = C.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipMalloc.restype = C.c_int
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipMemcpy.restype = C.c_int
    H.hipFree.argtypes = [C.c_void_p]
    H.hipFree.restype = C.c_int
    H.hipDeviceSynchronize.restype = C.c_int
    return H


class DeviceArray:
    def __init__(self, H, nbytes):
        self.H, self.p = H, C.c_void_p()
        assert H.hipMalloc(C.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        assert self.H.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, shape, dtype):
        out = np.empty(shape, dtype)
        assert out.nbytes == self.nbytes
        assert self.H.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.H.hipFree(self.p)


def noalloc_forward(ns, H, field):
    ny, nx = field.shape
    L = ns.load()
    d_in = DeviceArray(H, field.nbytes)
    d_out = DeviceArray(H, ny * (nx // 2 + 1) * 16)
    d_in.put(field.astype(np.float64))
    # leave the output dirty: the transform writes every element
    d_out.put(np.full((ny, nx // 2 + 1), 7.0 + 7.0j))
    err = C.create_string_buffer(256)
    rc = L.nsd_physical_to_spectral_noalloc(nx, ny, d_in.p, d_out.p, None, err, 256)
    assert rc == 0, err.value.decode()
    assert H.hipDeviceSynchronize() == 0
    out = d_out.get((ny, nx // 2 + 1), np.complex128)
    d_in.free()
    d_out.free()
    return out


@pytest.mark.parametrize("shapes", [[(256, 256), (64, 64)], [(128, 1024), (32, 16)]])
def test_noalloc_forward_matches_context_path(ns, shapes):
    H = hip_runtime(ns)
    for ny, nx in shapes:  # two calls in a row, the second on a smaller grid
        field = np.random.default_rng(nx * 7 + ny).standard_normal((ny, nx))
        s = ns.Spectral(nx, ny)
        s.init(taylor_green=False)
        ref = s.p2s(field)  # nsd_physical_to_spectral through the reference entry point
        got = noalloc_forward(ns, H, field)
        err = rel_l2(got, ref)
        print(f"{ny}x{nx}: rel L2 vs context path {err:.2e}", flush=True)
        assert err <= 1e-14
        assert np.all(got[s.mask == 0] == 0)
        s.close()


def test_noalloc_forward_rejects_grids_beyond_its_scratch(ns):
    H = hip_runtime(ns)
    L = ns.load()
    d = DeviceArray(H, 16)
    err = C.create_string_buffer(256)
    rc = L.nsd_physical_to_spectral_noalloc(2048, 2048, d.p, d.p, None, err, 256)
    assert rc == ns.SOLVER_INVALID_PARAM
    assert b"[16, 1024]" in err.value
    rc = L.nsd_physical_to_spectral_noalloc(64, 48, d.p, d.p, None, err, 256)
    assert rc == ns.SOLVER_INVALID_PARAM
    d.free()
