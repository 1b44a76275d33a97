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


@pytest.mark.parametrize("grid", [34, 38, 46, 58, 62, 74])
def test_grids_without_tile_edge_are_refused_with_the_reason(grid):
    """The deposit tiles the grid with edges in [4, 16]; a grid 2p with p a prime above 16 has
    none and is refused before any device work, with a message that says so."""
    from pbi4_amd.core import assign_workspace_bytes
    from pbi4_amd._lib import Pbi4Error
    with pytest.raises(Pbi4Error, match=rf"grid {grid} has no divisor in \[4, 16\]"):
        assign_workspace_bytes(grid, 4, True, 1000)
    assert assign_workspace_bytes(grid + 2, 4, True, 1000) > 1000 * 12   # 36, 40, 48, 60, 64, 76


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


def test_io_reader_dispatch_npz(tmp_path):
    # the host half of IO.read_catalogues: the npz reader hands the device front-end the raw
    # positions, the velocities and the (1 + z) / H(z) shift rule when multipoles are asked for
    # (the shift, wrap and normalisation themselves are checked bit-exactly on the GPU)
    import pbi4_amd as pb
    from pbi4_amd._lib import RSD_HZ, RSD_NONE
    from pbi4_amd.readers import read_catalogue
    rng = np.random.default_rng(0)
    pos = (rng.random((500, 3)) * 1000 - 10).astype(np.float32)      # some outside the box
    vel = (rng.standard_normal((500, 3)) * 300).astype(np.float32)
    np.savez(tmp_path / "cat.npz", pos=pos, vel=vel)
    io = pb.IO(config_file=_config(tmp_path))
    p = io.read_params()
    c = read_catalogue(p)
    assert c.rsd == RSD_HZ and c.npart == 500 and c.out_dtype is None
    assert np.array_equal(c.pos, pos) and np.array_equal(c.vel, vel)
    np.savez(tmp_path / "cat.npz", pos=pos.T)                           # (3, n), no velocities
    c = read_catalogue(p)
    assert c.rsd == RSD_NONE and c.vel is None and np.array_equal(c.pos, pos)


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


def test_script_command_line_and_config_write_back(tmp_path):
    # `import pbi4 as pb` + IO(scriptmode=True): command-line values override the JSON config,
    # derived parameters are added and the file is rewritten (pbi4/pbi4.py:26-77, 86-119)
    import os
    import subprocess
    import sys
    cfg = {"cat": "sancho", "box": 1000, "grid": 144.0, "interpolation": 4, "interlacing": True,
           "kbin": 2.0, "kcenter": 2.0, "iopen": True, "om": 0.3, "ol": 0.7, "z": 0.5,
           "multipoles": True, "los": [0.0, 0.0, 1.0], "cross": False, "config_file": "config.json"}
    (tmp_path / "config.json").write_text(json.dumps(cfg))
    script = ("import json, pbi4 as pb\n"
              "io = pb.IO(config_file='config.json', verbose=False, scriptmode=True)\n"
              "p = io.read_params()\n"
              "print(json.dumps({k: p[k] for k in ('grid', 'kbin', 'kcenter', 'interlacing', "
              "'nbins_ps', 'nbins_bisp', 'itot')}))\n"
              "print(json.dumps([float(x) for x in p['los']]))\n")
    (tmp_path / "run.py").write_text(script)
    env = dict(os.environ, PYTHONPATH=ROOT)
    out = subprocess.run([sys.executable, "run.py", "--kbin", "1.", "--kcenter", "1.", "--grid", "64",
                          "--los", "0", "2", "1", "--no-interlacing"], cwd=tmp_path, env=env,
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().splitlines()
    got = json.loads(lines[-2])
    assert got == {"grid": 64.0, "kbin": 1.0, "kcenter": 1.0, "interlacing": False,
                   "nbins_ps": 32, "nbins_bisp": 21, "itot": got["itot"]}
    assert np.allclose(json.loads(lines[-1]), [0.0, 0.4, 0.2])
    written = json.loads((tmp_path / "config.json").read_text())
    assert written["grid"] == 64.0 and written["interlacing"] is False
    assert written["itot"] == got["itot"] and len(written["kvec"]) == 32


def test_bench_finds_the_roofline_kernels_traffic():
    # the committed rocprofv3 traffic summary names the deposit kernel the bench reports
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", f"{ROOT}/bench.py")
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    t = b.pmc_traffic("k_deposit<4, 2, 20>")
    assert t is not None and t > 335544320          # at least the algorithmic bytes per launch


def test_assign_tile_edge_follows_particle_density():
    # the assignment workspace holds the tile regions, (T + 4)^3 cells per grid and tile: the
    # tile edge T is the largest divisor of N in [4, 16] with at most 8192 particles per tile on
    # average, but not below 8 while a divisor in [8, 16] exists (no GPU needed for the query)
    import ctypes
    from pbi4_amd import _lib
    L = _lib.lib()

    def ws(N, npart):
        d = _lib.AssignDesc(N, 4, 1, 1, 1000.0, (ctypes.c_float * 3)(0, 0, 0))
        return L.pbi4_assign_workspace(ctypes.byref(d), npart)

    def regions(N, T):
        return (N // T) ** 3 * 2 * (T + 4) ** 3 * 4

    def bases(N, T, npart):
        # with at most 8192 tiles the sort keeps each work-group's first slot per tile: one word
        # per (work-group, tile), min(2048, ceil(npart / 8192)) work-groups
        ntiles = (N // T) ** 3
        return 4 * max(1, min(2048, -(-npart // 8192))) * ntiles if ntiles <= 8192 else 0

    def rest(N, npart):                 # sort arrays and tile counters, below the regions
        return ws(N, npart) - 12 * npart

    for N, npart, T in [(128, 1_000_000, 16), (128, 1 << 24, 8), (256, 1 << 24, 16),
                        (192, 1 << 24, 12), (92, 60000, 4), (20, 1 << 20, 10)]:
        r = rest(N, npart) - bases(N, T, npart)
        assert regions(N, T) <= r <= regions(N, T) + 3 * 4 * ((N // T) ** 3 + 1) + 8192, (N, npart)


def test_import_sets_rocfft_kernel_build():
    import os
    # rocFFT builds its kernels with packed FP32 off, into the package's own kernel cache, and
    # takes none from the system cache (set on import, before torch is used)
    import pbi4_amd  # noqa: F401
    assert "-packed-fp32-ops" in os.environ["HIPRTC_COMPILE_OPTIONS_APPEND"]
    assert os.environ["ROCFFT_RTC_SYS_CACHE_PATH"].startswith("/nonexistent/")
    assert os.environ["ROCFFT_RTC_CACHE_PATH"] == os.environ["PBI4_ROCFFT_CACHE"]
    assert os.path.isdir(os.path.dirname(os.environ["ROCFFT_RTC_CACHE_PATH"]))
