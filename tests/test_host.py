"""CPU tests of the host side: library exports, drop-in configuration logic, triangle lists,
the synthetic catalogue generator and the serving admission policy (forced from the host)."""
import json
import re

import numpy as np
import pytest

from oracle import pbi4_oracle as O

ROOT = __import__("os").path.dirname(__import__("os").path.dirname(__import__("os").path.abspath(__file__)))


def test_library_loads_and_exports_every_declared_symbol():
    from pbi4_amd import _lib
    L = _lib.lib()                       # no compute call: there is no GPU here
    hdr = open(f"{ROOT}/include/pbi4_amd.h").read()
    declared = set(re.findall(r"^(?:int|size_t|const char \*)\s*\**(pbi4_\w+)\(", hdr, re.M))
    assert declared == set(_lib.SIGNATURES)
    for name in declared:
        assert getattr(L, name) is not None
    assert L.pbi4_version() >= 100


def test_library_errors_are_status_codes_without_gpu():
    import ctypes
    from pbi4_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    rc = L.pbi4_ctx_create(0, ctypes.byref(h))
    assert rc != 0 or h.value                # no GPU: an error code, never a crash
    if rc == 0:
        L.pbi4_ctx_destroy(h)
    # invalid descriptors are rejected before any device work
    d = _lib.AssignDesc(16, 7, 1, 0, 1.0, (ctypes.c_float * 3)(0, 0, 0))
    assert L.pbi4_assign_workspace(ctypes.byref(d), 100) == 0
    d = _lib.AssignDesc(34, 4, 1, 0, 1.0, (ctypes.c_float * 3)(0, 0, 0))   # 34 = 2 x 17
    assert L.pbi4_assign_workspace(ctypes.byref(d), 100) == 0
    d = _lib.AssignDesc(32, 4, 1, 0, 1.0, (ctypes.c_float * 3)(0, 0, 0))
    assert L.pbi4_assign_workspace(ctypes.byref(d), 100) > 100 * 12


def _config(tmp_path, **kw):
    cfg = {"cat": "npz", "box": 1000, "grid": 64.0, "interpolation": 4, "interlacing": True,
           "kbin": 1.0, "kcenter": 1.0, "iopen": True, "om": 0.3, "ol": 0.7, "z": 0.5,
           "multipoles": True, "los": [0.0, 0.0, 1.0], "cross": False,
           "cfile": str(tmp_path / "cfile.npz"), "npool": 4, "savedensity": False,
           "config_file": str(tmp_path / "config.json"), "ifile": str(tmp_path / "cat.npz"),
           "ps_outfile": str(tmp_path / "ps.dat"), "bs_outfile": str(tmp_path / "bs.dat")}
    cfg.update(kw)
    (tmp_path / "config.json").write_text(json.dumps(cfg))
    return str(tmp_path / "config.json")


def test_io_derived_params_and_config_rewrite(tmp_path):
    import pbi4_amd as pb
    io = pb.IO(config_file=_config(tmp_path))
    p = io.read_params()
    ref = O.derived_params(dict(p))
    for k in ("nbins_ps", "nbins_bisp", "itot", "kf", "kN", "Hz", "dk"):
        assert p[k] == ref[k], k
    assert np.array_equal(p["kvec"], ref["kvec"])
    assert (p["nbins_ps"], p["nbins_bisp"], p["itot"]) == (32, 21, 1156)
    # the config file is rewritten with the derived values (pbi4/pbi4.py:38-39)
    back = json.loads((tmp_path / "config.json").read_text())
    assert back["itot"] == 1156 and len(back["kvec"]) == 32


def test_io_read_catalogue_normalisation_and_rsd(tmp_path):
    import pbi4_amd as pb
    rng = np.random.default_rng(0)
    pos = (rng.random((500, 3)) * 1000 - 10).astype(np.float32)      # some outside the box
    vel = (rng.standard_normal((500, 3)) * 300).astype(np.float32)
    np.savez(tmp_path / "cat.npz", pos=pos, vel=vel)
    io = pb.IO(config_file=_config(tmp_path))
    p = io.read_params()
    out = io.read_catalogues()
    assert out.shape == (3, 500) and out.dtype == np.float32
    assert out.min() >= 0 and out.max() < 1
    # the reference's arithmetic (pbi4/pbi4.py:154, 241-243)
    s = (pos + vel * p["los"] / p["Hz"] * (1.0 + p["z"])).astype(np.float32).T
    s = (s + np.float32(1000)) % np.float32(1000)
    exp = np.minimum(1.0 - np.nextafter(0, 1), s / np.float32(1000))
    assert np.array_equal(out, exp.astype(np.float32))


def test_pipeline_itot_matches_reference_loop():
    from pbi4_amd.pipeline import Measurement, itot_count
    for nb, kb, kc, io in [(16, 2.0, 2.0, True), (21, 1.0, 1.0, True), (48, 1.0, 1.0, True),
                           (32, 2.0, 2.0, True), (10, 1.5, 0.5, False)]:
        assert itot_count(nb, kb, kc, io) == O.itot_count(nb, kb, kc, io)
        assert itot_count(nb, kb, kc, io, True) == O.itot_count(nb, kb, kc, io, True)
    m = Measurement(grid=256, kbin=2.0, kcenter=2.0, nbins_bisp=16)
    assert m.itot() == 564 and m.nb_ps() == 64
    assert Measurement(grid=128, kbin=2.0, kcenter=2.0, nbins_bisp=0).nb_bisp() == 21


def test_zeldovich_catalogue_cpu():
    from pbi4_amd.catalogue import growth_rate, hubble, rsd_factor, zeldovich
    pos, vel = zeldovich(16, box=1000.0, seed=3)
    pos2, _ = zeldovich(16, box=1000.0, seed=3)
    assert pos.shape == (3, 4096) and vel.shape == (3, 4096)
    assert bool((pos == pos2).all())                          # seeded
    assert float(pos.min()) >= 0.0 and float(pos.max()) < 1000.0
    assert abs(hubble() - 130.86252328302402) < 1e-9         # pbi4.py:102 with z=0.5
    assert abs(rsd_factor() - 1.5 / hubble()) < 1e-15
    assert 0.7 < growth_rate() < 0.8
    # displacement rms of the lattice: finite and well below the box
    q = (np.arange(16) + 0.5) * (1000.0 / 16)
    d = pos[2].numpy().reshape(16, 16, 16) - q[None, None, :]
    d = (d + 500) % 1000 - 500
    assert 1.0 < d.std() < 200.0


def test_admission_policy():
    from pbi4_amd.serving import ADMIT, DEFER, REJECT, WAIT, admission
    GB = 1 << 30
    assert admission(1.0, 2.0, True, GB, 10 * GB, 0) == WAIT          # not arrived yet
    assert admission(2.0, 1.0, False, GB, 10 * GB, 4) == DEFER         # no slot
    assert admission(2.0, 1.0, True, 12 * GB, 10 * GB, 2) == DEFER     # memory held by others
    assert admission(2.0, 1.0, True, 12 * GB, 10 * GB, 0) == REJECT    # never fits
    assert admission(2.0, 1.0, True, 3 * GB, 10 * GB, 3) == ADMIT


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, f"{ROOT}/bench.py", "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
