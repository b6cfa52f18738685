"""Write the golden vectors under tests/golden/ from the reference's own code.

Run in the development container (where /root/reference exists) after
`make -C oracle ref`; it executes only oracle/_ref/libref_forcing.so and
oracle/_ref/libref_config.so, compiled unmodified from the reference's
src/common/forcing.c and src/common/config_parser.c.

Outputs:
  forcing_golden.npz   ns_forcing_generate draws (seed 42, shell [4, 6], grids
                       16x16, 32x32, 64x64, three calls, white and coloured)
                       plus the 4x4 case of test_solvers.c:239-295
  config_golden.json   keys/values the reference parser produces for
                       parameters.ini + config_override.json, and the
                       validation result of the schema of test_solvers.c:327-343
parameters.ini and config_override.json are the reference's own fixture
files (config/parameters.ini, tests/config_override.json), copied as data.
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402


def k2_grid(n, L=2 * np.pi):
    nkx = n // 2 + 1
    kx = 2 * np.pi * np.arange(nkx) / L
    j = np.arange(n)
    ky = 2 * np.pi * np.where(j <= n // 2, j, j - n) / L
    return (kx[None, :] ** 2 + ky[:, None] ** 2), nkx


def forcing_vectors(ref):
    out = {}
    for n in (16, 32, 64):
        k2, nkx = k2_grid(n)
        for colored in (0, 1):
            f = O.OracleForcing(k2, nkx, n, 4.0, 6.0, energy_target=1.5, seed=42,
                                colored=bool(colored), correlation_time=0.2, library=ref)
            assert f.rc == 0
            draws = [f.generate(0.05) for _ in range(3)]
            out[f"n{n}_c{colored}"] = np.stack(draws)
            out[f"n{n}_c{colored}_rng_state"] = np.array([f.state.rng_state], dtype=np.uint64)
            f.close()
    # test_solvers.c:239-295 (4x4, k2 = i^2 + j^2, shell [0.5, 2.5], seed 42)
    k2 = np.array([[i * i + j * j for i in range(4)] for j in range(4)], dtype=np.float64)
    for colored, dt, tau in ((0, 0.1, 0.5), (1, 0.05, 0.2)):
        f = O.OracleForcing(k2, 4, 4, 0.5, 2.5, energy_target=1.5, seed=42,
                            colored=bool(colored), correlation_time=tau, library=ref)
        out[f"t4_c{colored}"] = np.stack([f.generate(dt) for _ in range(2)])
        f.close()
    return out


def config_vectors(path):
    L = C.CDLL(path)

    class Cfg(C.Structure):
        _fields_ = [("entries", C.c_void_p), ("count", C.c_size_t), ("capacity", C.c_size_t)]

    class Entry(C.Structure):
        _fields_ = [("key", C.c_char_p), ("value", C.c_char_p)]

    cfg = Cfg()
    L.ns_config_init(C.byref(cfg))
    L.ns_config_set.argtypes = [C.POINTER(Cfg), C.c_char_p, C.c_char_p]
    for k, v in (("solver.type", "spectral"), ("grid.nx", "64"), ("grid.ny", "64"),
                 ("time.dt", "0.005")):
        L.ns_config_set(C.byref(cfg), k.encode(), v.encode())
    files = (C.c_char_p * 2)(os.path.join(HERE, "parameters.ini").encode(),
                             os.path.join(HERE, "config_override.json").encode())
    err = C.create_string_buffer(256)
    L.ns_config_load_files.argtypes = [C.POINTER(Cfg), C.POINTER(C.c_char_p), C.c_size_t,
                                       C.c_char_p, C.c_size_t]
    rc = L.ns_config_load_files(C.byref(cfg), files, 2, err, 256)
    entries = C.cast(cfg.entries, C.POINTER(Entry))
    kv = [[entries[i].key.decode(), entries[i].value.decode()] for i in range(cfg.count)]
    L.ns_config_free(C.byref(cfg))
    return {"load_rc": rc, "entries": kv}


def main():
    ref_forcing = O.ref_forcing_lib()
    if ref_forcing is None or not os.path.exists(O.REF_CONFIG):
        sys.exit("oracle/_ref missing: run `make -C oracle ref` first")
    np.savez_compressed(os.path.join(HERE, "forcing_golden.npz"), **forcing_vectors(ref_forcing))
    with open(os.path.join(HERE, "config_golden.json"), "w") as fh:
        json.dump(config_vectors(O.REF_CONFIG), fh, indent=1)
    print("golden vectors written")


if __name__ == "__main__":
    main()
