"""CPU-side checks of the product library: it loads, exports every symbol the
headers in include/ declare, and its host-only parts (ns_config parser,
time-step helpers) behave like the reference's."""
import ctypes as C
import json
import os
import re

import pytest

from conftest import GOLDEN, ROOT


def declared_symbols():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w \*]*?\b([a-z_][a-z0-9_]*)\s*\(", text,
                             flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "sizeof", "return"):
                names.add(name)
    return names


def test_library_exports_every_declared_symbol(ns):
    L = ns.load()
    decl = declared_symbols()
    assert len(decl) >= 80
    missing = [n for n in sorted(decl) if not hasattr(L, n)]
    assert not missing, missing
    assert set(ns.EXPORTED) == decl
    # a 64-bit context pointer passed without argtypes is truncated to a C int
    untyped = [n for n in sorted(decl) if getattr(L, n).argtypes is None]
    assert not untyped, untyped


def test_nonpow2_or_missing_device_fails_loudly(ns):
    L = ns.load()
    # reference allocate returns NULL on failure; here also without a HIP device
    p = L.ns_spectral_allocate(48, 48, 6.28, 6.28, 100.0)
    assert not p


def load_ref_fixture(ns, L):
    cfg = ns.NSConfig()
    L.ns_config_init(C.byref(cfg))
    for k, v in (("solver.type", "spectral"), ("grid.nx", "64"), ("grid.ny", "64"),
                 ("time.dt", "0.005")):
        assert L.ns_config_set(C.byref(cfg), k.encode(), v.encode()) == 0
    files = (C.c_char_p * 2)(os.path.join(GOLDEN, "parameters.ini").encode(),
                             os.path.join(GOLDEN, "config_override.json").encode())
    err = C.create_string_buffer(256)
    rc = L.ns_config_load_files(C.byref(cfg), files, 2, err, 256)
    return cfg, rc, err


def entries(cfg):
    class Entry(C.Structure):
        _fields_ = [("key", C.c_char_p), ("value", C.c_char_p)]
    e = C.cast(cfg.entries, C.POINTER(Entry))
    return [[e[i].key.decode(), e[i].value.decode()] for i in range(cfg.count)]


def test_config_matches_reference_golden(ns):
    # test_solvers.c:297-347 sequence on the reference's own fixture files
    L = ns.load()
    cfg, rc, err = load_ref_fixture(ns, L)
    gold = json.load(open(os.path.join(GOLDEN, "config_golden.json")))
    assert rc == gold["load_rc"] == 0, err.value
    assert entries(cfg) == gold["entries"]
    g = lambda k: L.ns_config_get_string(C.byref(cfg), k.encode(), b"").decode()
    assert g("solver.type") == "finite_difference"
    assert L.ns_config_get_int(C.byref(cfg), b"grid.nx", 0) == 192
    assert L.ns_config_get_double(C.byref(cfg), b"time.dt", 0.0) == pytest.approx(0.001)
    assert L.ns_config_get_double(C.byref(cfg), b"forcing.energy_target", 0.0) == 1.25
    S = ns.NSConfigSchemaEntry
    solver_allowed = (C.c_char_p * 2)(b"spectral", b"finite_difference")
    forcing_allowed = (C.c_char_p * 3)(b"none", b"white", b"colored")
    D, I, ST = 2, 1, 0
    schema = (S * 12)(
        S(b"solver.type", ST, 1, 0, 0, 0, 0, solver_allowed, 2),
        S(b"grid.nx", I, 1, 2.0, 0, 1, 0, None, 0), S(b"grid.ny", I, 1, 2.0, 0, 1, 0, None, 0),
        S(b"domain.Lx", D, 1, 1e-6, 0, 1, 0, None, 0), S(b"domain.Ly", D, 1, 1e-6, 0, 1, 0, None, 0),
        S(b"fluid.Re", D, 1, 1e-6, 0, 1, 0, None, 0), S(b"time.dt", D, 1, 1e-9, 0, 1, 0, None, 0),
        S(b"time.t_end", D, 1, 1e-9, 0, 1, 0, None, 0),
        S(b"forcing.energy_target", D, 0, 0, 0, 1, 0, None, 0),
        S(b"forcing.type", ST, 0, 0, 0, 0, 0, forcing_allowed, 3),
        S(b"spectral.hyperviscosity.coefficient", D, 0, 0, 0, 1, 0, None, 0),
        S(b"spectral.hyperviscosity.order", I, 0, 0, 0, 1, 0, None, 0))
    assert L.ns_config_validate(C.byref(cfg), schema, 12, err, 256) == 0, err.value
    bad = (S * 1)(S(b"grid.nx", I, 1, 500.0, 0, 1, 0, None, 0))
    assert L.ns_config_validate(C.byref(cfg), bad, 1, err, 256) == -1
    assert b"below minimum" in err.value
    L.ns_config_free(C.byref(cfg))


REF_CONFIG = os.path.join(ROOT, "oracle", "_ref", "libref_config.so")


@pytest.mark.skipif(not os.path.exists(REF_CONFIG), reason="reference config build absent")
def test_config_parser_vs_live_reference_build(ns, tmp_path):
    L = ns.load()
    R = C.CDLL(REF_CONFIG)
    ini = tmp_path / "a.ini"
    ini.write_text("# c\n; c\ntop = 1\n[ s1 ]\n k = v v \nempty=\n[s2.sub]\nx = 3.5e2\n")
    js = tmp_path / "b.JSON"
    js.write_text('{"a": {"b": {"c": "s\\"q\\\\\\/\\t"}, "n": -1.25e-3, "t": true,'
                  ' "f": false, "z": null}, "top": 7, "e": {}}')
    for L_ in (L, R):
        L_.ns_config_set.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p]
        L_.ns_config_load_files.argtypes = [C.c_void_p, C.POINTER(C.c_char_p), C.c_size_t,
                                            C.c_char_p, C.c_size_t]
    out = []
    for L_ in (L, R):
        cfg = ns.NSConfig()
        L_.ns_config_init(C.byref(cfg))
        files = (C.c_char_p * 2)(str(ini).encode(), str(js).encode())
        err = C.create_string_buffer(256)
        rc = L_.ns_config_load_files(C.byref(cfg), files, 2, err, 256)
        out.append((rc, entries(cfg)))
        L_.ns_config_free(C.byref(cfg))
    assert out[0] == out[1]
    for text in ('{"a": [1]}', '{"a": 1} x', '{"a" 1}', "[sec\nk=v\n", "novalue\n"):
        res = []
        for L_ in (L, R):
            p = tmp_path / ("c.json" if text.startswith("{") else "c.ini")
            p.write_text(text)
            cfg = ns.NSConfig()
            L_.ns_config_init(C.byref(cfg))
            err = C.create_string_buffer(256)
            files = (C.c_char_p * 1)(str(p).encode())
            res.append(L_.ns_config_load_files(C.byref(cfg), files, 1, err, 256))
            L_.ns_config_free(C.byref(cfg))
        assert res[0] == res[1] == -1


def test_time_integration_helpers(ns):
    # api_reference.md:178-193 (declared, never defined by the reference)
    L = ns.load()
    assert L.compute_viscous_timestep(0.1, 0.2, 50.0) == pytest.approx(0.5 * 50 * 0.02 / 0.3)
    assert L.compute_cfl_condition(2.0, 1.0, 0.1, 0.1, 0.5) == pytest.approx(0.5 / 30.0)
    p = ns.AdaptiveTimeParams(0.5, 1e-4, 1e-1, 0.9, 1)
    dt = L.adaptive_timestep(10.0, 0.01, 0.01, 1000.0, C.byref(p))
    assert dt == pytest.approx(0.9 * min(0.5 * 0.01 / 10.0, 0.5 * 1000 * 1e-4 / 0.02))
    assert L.adaptive_timestep(1e9, 0.01, 0.01, 1000.0, C.byref(p)) == pytest.approx(1e-4)
