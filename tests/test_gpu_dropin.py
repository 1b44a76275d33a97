"""GPU tests of the drop-in layer: the reference's script call sequence, the f2py-named
functions, the reference-layout array views, the fused P+B entry point and the serving loop
(each result the same alone and in a mixed batch).  Tolerances as in test_gpu_parity."""
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import pbi4_oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _clustered(n, seed, box=1000.0):
    rng = np.random.default_rng(seed)
    c = rng.random((6, 3)) * box
    k = rng.integers(0, 6, n)
    p = c[k] + 50.0 * rng.standard_normal((n, 3))
    v = 300.0 * rng.standard_normal((n, 3))
    return p.astype(np.float32), v.astype(np.float32)


def _second_tracer(pos, seed, jitter=0.004):
    """A second catalogue of the same clustering: a random 3/4 of the particles, each moved by
    a small random offset (in box units), as two tracers of one field.  Positions (3, n) in
    [0, 1)."""
    rng = np.random.default_rng(seed)
    keep = rng.random(pos.shape[1]) < 0.75
    q = pos[:, keep] + jitter * rng.standard_normal((3, int(keep.sum())))
    return (q % 1.0).astype(pos.dtype)


def _pk_close(mine, ref, rows):
    ok = ref[rows[-1]] > 0
    for r in rows[1:-1]:
        assert np.all(np.abs(mine[r][ok] - ref[r][ok]) <= 1e-4 * np.abs(ref[1][ok])), r


def _ulp32(x):
    """Spacing of float32 at |x|: the resolution of a value stored in the reference's float32
    output records (pbi4/pbi4.py:482-534)."""
    return np.spacing(np.abs(np.asarray(x, dtype=np.float32))).astype(np.float64)


def _bk_close(mine, ref, rtol=1e-4):
    """Both sides are float32 records.  B0 is read back as (B0 + BSN) - BSN, so besides rtol of
    the bin's scale the difference may hold the float32 rounding of the two stored values."""
    gauss = np.sqrt(np.abs(ref[3] * ref[4] * ref[5]) / np.maximum(ref[8], 1))
    b0 = np.abs(ref[6] - ref[7])
    for r in (0, 1, 2, 8):
        assert np.array_equal(mine[r], ref[r]), r
    for r in (3, 4, 5, 7):
        assert np.allclose(mine[r], ref[r], rtol=rtol), r
    for r in [6] + ([9, 10] if ref.shape[0] == 11 else []):
        bref = ref[r] - (ref[7] if r == 6 else 0)
        bm = mine[r] - (mine[7] if r == 6 else 0)
        store = _ulp32(ref[r]) + (_ulp32(ref[7]) if r == 6 else 0)
        err = np.maximum(np.abs(bm - bref) - store, 0) / np.maximum(np.maximum(np.abs(bref), gauss), b0)
        assert err.max() <= rtol, (r, err.max())


# ------------------------------------------------------------------ the script sequence


@pytest.mark.parametrize("multipoles,precision", [(True, "fp32"), (True, "bf16x3"), (False, "fp32")])
def test_script_sequence(tmp_path, multipoles, precision):
    import pbi4_amd as pb
    pos, vel = _clustered(60000, 11)
    np.savez(tmp_path / "cat.npz", pos=pos, vel=vel)
    cfg = {"cat": "npz", "box": 1000, "grid": 32.0, "interpolation": 4, "interlacing": True,
           "kbin": 1.0, "kcenter": 1.0, "iopen": True, "om": 0.3, "ol": 0.7, "z": 0.5,
           "multipoles": multipoles, "los": [0.0, 0.0, 1.0], "cross": False,
           "cfile": str(tmp_path / "cfile.npz"), "npool": 4, "savedensity": False,
           "config_file": str(tmp_path / "config.json"), "ifile": str(tmp_path / "cat.npz"),
           "ps_outfile": str(tmp_path / "ps.dat"), "bs_outfile": str(tmp_path / "bs.dat"),
           "precision": precision}
    (tmp_path / "config.json").write_text(json.dumps(cfg))
    # Tutorials/Script/test.py, line by line
    io = pb.IO(config_file=str(tmp_path / "config.json"), verbose=True, scriptmode=False)
    params = io.read_params()
    pos_n = io.read_catalogues()
    density = pb.DENSITIES(params)
    npart, dcl = density.assign_grid(pos_n)
    dcl = density.compute_density(npart, dcl)
    ms = pb.MEASUREMENTS(params)
    if multipoles:
        dcl2, dcl4 = density.compute_density_multipoles(dcl)
        ps = ms.powerspectrum(npart, dcl, dcl2, dcl4)
    else:
        dcl2 = dcl4 = None
        ps = ms.powerspectrum(npart, dcl)
    counts = ms.triangle_counts()
    bs = ms.bispectrum(npart, counts, dcl, dcl2, dcl4)
    ms.save_powerspectrum(ps)
    ms.save_bispectrum(bs)
    ms.save_triangle_counts(counts)
    assert os.path.getsize(params["ps_outfile"]) > 0 and os.path.getsize(params["bs_outfile"]) > 0

    # the reference algorithm on the same normalised positions (FP64 oracle)
    N = 32
    g = O.assign_double_grid(N, 4, npart, pos_n.astype(np.float64))
    d = O.compute_density(npart, g, N, 4, True, 1000.0)
    nb_ps, nb_b, itot = params["nbins_ps"], params["nbins_bisp"], params["itot"]
    rc = O.count_triangles_exact(nb_b, N, 1.0, 1.0, True).astype(np.float64) * N ** 3
    assert np.array_equal(counts, rc)
    kf = np.float32(params["kf"])
    if multipoles:
        d2, d4 = O.compute_density_multipoles(d, N, 1000.0, params["los"])
        pref = O.measure_pk_multipoles(N, 1000.0, 1.0, 1.0, nb_ps, d, d2, d4)
        bref = O.measure_bk(nb_b, N, 1.0, 1.0, True, kf, itot, npart, rc, d, d2, d4)
        mine_p = np.array([ps["avgk"], ps["P0+PSN"], ps["P2"], ps["P4"], ps["Nmodes"]], dtype=np.float64)
        assert np.array_equal(mine_p[4], pref[4].astype(np.float32))
        _pk_close(mine_p, pref.astype(np.float32).astype(np.float64), [0, 1, 2, 3, 4])
        names = ["k1", "k2", "k3", "Pk1", "Pk2", "Pk3", "B0+BSN", "BSN", "Ntr", "B2", "B4"]
    else:
        pref = O.measure_pk(N, 1000.0, 1.0, 1.0, nb_ps, d, d)
        bref = O.measure_bk(nb_b, N, 1.0, 1.0, True, kf, itot, npart, rc, d)
        mine_p = np.array([ps["avgk"], ps["P0+PSN"], ps["Nmodes"]], dtype=np.float64)
        assert np.array_equal(mine_p[2], pref[2].astype(np.float32))
        _pk_close(mine_p, pref.astype(np.float32).astype(np.float64), [0, 1, 2])
        names = ["k1", "k2", "k3", "Pk1", "Pk2", "Pk3", "B0+BSN", "BSN", "Ntr"]
    mine_b = np.array([bs[n] for n in names], dtype=np.float64)
    _bk_close(mine_b, bref.astype(np.float32).astype(np.float64))
    # the cached counts file is read back on the next call (pbi4/pbi4.py:466-470)
    assert np.array_equal(ms.triangle_counts(), counts)


def test_cross_bispectrum_measurements(tmp_path):
    # MEASUREMENTS with params["cross"]: counts over the cross loop, B of two catalogues
    # (pbi4/pbi4.py:472-496) and the five-column bs file (:579-582)
    import pbi4_amd as pb
    N, n, nb = 32, 40000, 6
    pa = (_clustered(n, 12)[0].T % 1000.0 / 1000.0).astype(np.float32)
    pb_ = _second_tracer(pa, 13)
    nb_part = pb_.shape[1]
    itc = O.itot_count(nb, 1.0, 1.0, True, True)
    params = {"grid": N, "interpolation": 4, "interlacing": True, "box": 1000.0,
              "kf": 2 * np.pi / 1000, "kN": np.pi / 1000 * N, "multipoles": False,
              "los": np.array([0.0, 0.0, 1.0]), "cross": True, "nbins_bisp": nb, "kbin": 1.0,
              "kcenter": 1.0, "iopen": True, "itot": itc, "npool": 4,
              "cfile": str(tmp_path / "cfile.npz"), "bs_outfile": str(tmp_path / "bs.dat")}
    den = pb.DENSITIES(params)
    _, ga = den.assign_grid(pa)
    _, gb = den.assign_grid(pb_)
    da = den.compute_density(n, ga)
    db = den.compute_density(nb_part, gb)
    ms = pb.MEASUREMENTS(params)
    counts = ms.triangle_counts()
    bs = ms.bispectrum(n, counts, da, db)
    assert bs.dtype.names == ("k1", "k2", "k3", "B0", "Ntr") and len(bs) == itc
    ms.save_bispectrum(bs)
    cols = np.loadtxt(params["bs_outfile"])
    assert cols.shape == (itc, 5)
    oa = O.compute_density(n, O.assign_double_grid(N, 4, n, pa.astype(np.float64)), N, 4, True, 1000.0)
    ob = O.compute_density(nb_part, O.assign_double_grid(N, 4, nb_part, pb_.astype(np.float64)), N, 4,
                           True, 1000.0)
    ref = O.measure_bk_cross(nb, N, 1.0, 1.0, True, np.float32(params["kf"]), itc, n, counts, oa, ob)
    mine = np.array([bs[c] for c in ("k1", "k2", "k3", "B0", "Ntr")], dtype=np.float64)
    _bk_cross_close(mine, ref.astype(np.float32).astype(np.float64), oa, ob, N, nb,
                    np.float32(params["kf"]))


def test_cross_power_spectrum():
    import pbi4_amd.powerI4 as P
    N = 32
    pa, _ = _clustered(30000, 1)
    pb_, _ = _clustered(30000, 2)
    a = O.compute_density(30000, O.assign_double_grid(N, 3, 30000, (pa.T % 1000.0 / 1000).astype(np.float64)), N, 3, 1, 1000.0)
    b = O.compute_density(30000, O.assign_double_grid(N, 3, 30000, (pb_.T % 1000.0 / 1000).astype(np.float64)), N, 3, 1, 1000.0)
    mine = P.measure_pk(N, 1000.0, 1.0, 1.0, 16, a, b)
    ref = O.measure_pk(N, 1000.0, 1.0, 1.0, 16, a, b)
    assert np.array_equal(mine[2], ref[2])
    scale = np.sqrt(O.measure_pk(N, 1000.0, 1.0, 1.0, 16, a, a)[1] * O.measure_pk(N, 1000.0, 1.0, 1.0, 16, b, b)[1])
    assert np.all(np.abs(mine[1] - ref[1]) <= 1e-4 * scale)


# ------------------------------------------------------------------ f2py-named functions


def test_f2py_power_functions():
    import pbi4_amd.powerI4 as P
    N, n = 16, 20000
    pos = (_clustered(n, 4)[0].T % 1000.0 / 1000.0).astype(np.float32)
    p64 = pos.astype(np.float64)
    for order in (1, 2, 3, 4):
        g = P.assign_double_grid(N, order, n, pos, np.ones(n))
        r = O.assign_double_grid(N, order, n, p64)
        assert g.shape == (N, N, N) and g.flags.f_contiguous
        assert np.abs(g - r).max() <= 2e-6 * np.abs(r).max()
        g = P.assign_single_grid(N, order, n, pos, np.ones(n))
        r = O.assign_single_grid(N, order, n, p64)
        assert np.abs(g - r).max() <= 2e-6 * np.abs(r).max()
    F = np.fft.fftn(O.assign_double_grid(N, 4, n, p64) / n)
    for inter in (True, False):
        mine = P.fcomb(N, 4, inter, 1000.0, F)
        ref = O.fcomb(N, 4, inter, 1000.0, F)
        assert np.abs(mine - ref).max() <= 2e-6 * np.abs(ref).max(), inter
    d = O.fcomb(N, 4, True, 1000.0, F)
    assert np.allclose(P.compute_q2k(N, 1000.0, d, 3, 1), O.compute_q2k(N, 1000.0, d, 3, 1),
                       atol=2e-6 * np.abs(d).max())
    assert np.allclose(P.compute_q4k(N, 1000.0, d, 3, 3, 2, 1), O.compute_q4k(N, 1000.0, d, 3, 3, 2, 1),
                       atol=2e-6 * np.abs(d).max())
    d2, d4 = O.compute_density_multipoles(d, N, 1000.0, [0.0, 0.0, 1.0])
    mine = P.measure_pk_multipoles(N, 1000.0, 2.0, 2.0, 4, d, d2, d4)
    ref = O.measure_pk_multipoles(N, 1000.0, 2.0, 2.0, 4, d, d2, d4)
    assert np.array_equal(mine[4], ref[4])
    _pk_close(mine, ref, [0, 1, 2, 3, 4])
    mine = P.measure_pk(N, 1000.0, 1.0, 1.0, 8, d, d)
    ref = O.measure_pk(N, 1000.0, 1.0, 1.0, 8, d, d)
    assert np.array_equal(mine[2], ref[2])
    _pk_close(mine, ref, [0, 1, 2])


@pytest.mark.parametrize("L,G", [(16, 8), (12, 12), (16, 10), (9, 5), (64, 24)])
def test_reduce_array_bit_exact(L, G):
    # the device copy in FP64 against the oracle restatement (pinned on the CPU to the
    # reference's own reduce_array)
    import pbi4_amd.powerI4 as P
    rng = np.random.default_rng(3 + L + G)
    din = np.asfortranarray(rng.standard_normal((L, L, L)) + 1j * rng.standard_normal((L, L, L)))
    mine = P.reduce_array(L, din, G)
    assert mine.shape == (G, G, G) and mine.dtype == np.complex128
    assert np.array_equal(mine, O.reduce_array(L, din, G))


@pytest.mark.parametrize("pdt,vdt,rsd,obj,box", [
    (np.float32, np.float32, 0, False, 1000), (np.float32, np.float32, 1, False, 1000),
    (np.float32, np.float32, 2, False, 2500.5), (np.float64, np.float32, 1, False, 1000),
    (np.float64, np.float64, 0, False, 777.7), (np.float32, np.float32, 1, True, 1000),
    (np.float32, np.float32, 0, True, 1000)])
def test_catalogue_positions_bit_exact(pdt, vdt, rsd, obj, box):
    # the device shift / wrap / normalisation against the reference's numpy lines, including
    # positions far outside the box and the edges 0, -0, box, -box
    from pbi4_amd.core import default_context
    ctx = default_context()
    rng = np.random.default_rng(int(box) + 10 * rsd + obj)
    n = 20000
    pos = ((rng.random((n, 3)) * 3.4 - 1.2) * box).astype(pdt)
    pos[:8] = np.array([0.0, -0.0, box, -box, 2 * box, -2.5 * box, box * (1 - 1e-7), 1e-30])[:, None]
    vel = (rng.standard_normal((n, 3)) * 400).astype(vdt)
    los = np.array([0.0, 2.0, 1.0]) / 5.0
    hz, z = 130.86252328302402, 0.5
    ref = O.catalogue_positions(pos, vel, rsd, los, hz, z, box, object_rows=obj)
    # the object-array rows hold Python floats: float64 arithmetic throughout
    out_dtype = torch.float64 if obj or pdt == np.float64 else torch.float32
    dev = ctx.catalogue_positions(torch.from_numpy(pos).cuda(), torch.from_numpy(vel).cuda() if rsd else None,
                                  rsd, tuple(los), hz, 1.0 + z, float(box), out_dtype=out_dtype)
    assert np.array_equal(dev.cpu().numpy().astype(np.float64), np.asarray(ref, dtype=np.float64))


def test_read_catalogues_gadget_and_quijote(tmp_path):
    # IO.read_catalogues end to end: host reader, device front-end, DeviceCatalogue as the
    # reference's float64 (3, n) array; then the P(k) of that catalogue
    import pbi4_amd as pb
    from pbi4_amd import readers as R
    from test_readers import _fof_file, _gadget_file
    rng = np.random.default_rng(8)
    n = 30000
    p_all = (rng.random((n + 5, 3)) * 1000).astype(np.float32)
    v_all = (rng.standard_normal((n + 5, 3)) * 300).astype(np.float32)
    _gadget_file(str(tmp_path / "snap"), 2, "<", [5, n, 0, 0, 0, 0], p_all, v_all, time=2 / 3, redshift=0.5)
    cfg = {"cat": "gadget", "ifile": str(tmp_path / "snap"), "box": 1000, "grid": 32.0,
           "interpolation": 3, "interlacing": True, "kbin": 1.0, "kcenter": 1.0, "iopen": True,
           "om": 0.3, "ol": 0.7, "z": 0.5, "multipoles": True, "los": [0.0, 1.0, 1.0],
           "cross": False, "config_file": str(tmp_path / "c.json")}
    (tmp_path / "c.json").write_text(json.dumps(cfg))
    io = pb.IO(config_file=str(tmp_path / "c.json"))
    params = io.read_params()
    cat = io.read_catalogues()
    p, v = R.read_gadget(str(tmp_path / "snap"), 1, velocities=True)
    ref = O.catalogue_positions(p, v, 1, params["los"], params["Hz"], params["z"], params["box"])
    assert np.asarray(cat).dtype == np.float64 and np.array_equal(np.asarray(cat), ref)
    npart, g = pb.DENSITIES(params).assign_grid(cat)
    assert npart == n
    # Quijote FoF halos with a redshift-space shift: float64 positions (object-array branch)
    d = tmp_path / "groups_003"
    d.mkdir()
    _fof_file(str(d / "group_tab_003.0"), rng.random(n).astype(np.float32), p_all[:n] * 1e3, v_all[:n], n, 1)
    cfg.update({"cat": "quijote", "ifile": str(tmp_path)})
    (tmp_path / "c.json").write_text(json.dumps(cfg))
    io = pb.IO(config_file=str(tmp_path / "c.json"))
    params = io.read_params()
    cat = io.read_catalogues()
    assert cat.pos.dtype == torch.float64
    _, gp, gv = R.read_fof_groups(str(tmp_path), 3)
    ref = O.catalogue_positions(gp / np.float32(1e3), gv * np.float32(1.5), 1, params["los"],
                                params["Hz"], params["z"], params["box"], object_rows=True)
    assert np.array_equal(np.asarray(cat), ref)


def test_density_file_round_trip(tmp_path):
    # DENSITIES.save_density / read_density (pbi4/pbi4.py:257-288, :390-396): the same grid reads
    # back the identical device field; a smaller grid is the device reduce_array of the saved
    # array, whose P(k) matches the FP64 oracle on the oracle's reduction of that array
    import pbi4_amd as pb
    L, N, n = 32, 16, 50000
    pos = (_clustered(n, 21)[0].T % 1000.0 / 1000.0).astype(np.float32)
    base = {"box": 1000.0, "interpolation": 4, "interlacing": True, "kbin": 1.0, "kcenter": 1.0,
            "iopen": True, "multipoles": False, "los": np.array([0.0, 0.0, 1.0]), "cross": False,
            "dfile": str(tmp_path / "density.npz")}
    pL = {**base, "grid": L, "kf": 2 * np.pi / 1000.0, "kN": np.pi / 1000.0 * L}
    dens = pb.DENSITIES(pL)
    npart, g = dens.assign_grid(pos)
    d = dens.compute_density(npart, g)
    dens.save_density(d, npart, pL)
    with np.load(pL["dfile"], allow_pickle=False) as f:
        assert np.array_equal(f["header"], [L, npart, pL["kf"], pL["kN"]])
        saved = np.asarray(f["data"])
    npart2, d_same = dens.read_density(pL)
    assert npart2 == npart and torch.equal(d_same.delta, d.delta)
    pN = {**base, "grid": N, "kf": 2 * np.pi / 1000.0, "kN": np.pi / 1000.0 * N}
    pN.update({"nbins_ps": N // 2, "kvec": np.arange(1, N // 2 + 1) * pN["kf"]})
    _, d_red = pb.DENSITIES(pN).read_density(pN)
    red = O.reduce_array(L, saved, N)
    assert np.array_equal(np.asarray(d_red), red.astype(np.complex64).astype(np.complex128))
    ps = pb.MEASUREMENTS(pN).powerspectrum(npart, d_red)
    ref = O.measure_pk(N, 1000.0, 1.0, 1.0, N // 2, red, red)
    mine = np.array([ps["avgk"], ps["P0+PSN"], ps["Nmodes"]], dtype=np.float64)
    assert np.array_equal(mine[2], ref[2].astype(np.float32))
    _pk_close(mine, ref.astype(np.float32).astype(np.float64), [0, 1, 2])
    with pytest.raises(SystemExit):
        pb.DENSITIES({**base, "grid": 2 * L}).read_density()


def test_f2py_bispectrum_functions():
    import pbi4_amd.bispectrumI4 as B
    N, n, nb = 32, 40000, 6
    pos = (_clustered(n, 9)[0].T % 1000.0 / 1000.0).astype(np.float64)
    d = O.compute_density(n, O.assign_double_grid(N, 4, n, pos), N, 4, True, 1000.0)
    itot = O.itot_count(nb, 1.0, 1.0, True)
    counts = B.count_triangles(nb, N, 1.0, 1.0, True, False, 4, itot)
    assert np.array_equal(counts, O.count_triangles_exact(nb, N, 1.0, 1.0, True) * float(N) ** 3)
    kf = np.float32(2 * np.pi / 1000.0)
    _bk_close(B.measure_bk(nb, N, 1.0, 1.0, True, kf, 4, itot, n, counts, d),
              O.measure_bk(nb, N, 1.0, 1.0, True, kf, itot, n, counts, d))
    d2, d4 = O.compute_density_multipoles(d, N, 1000.0, [0.0, 1.0, 0.0])
    _bk_close(B.measure_bk_multipoles(nb, N, 1.0, 1.0, True, kf, 4, itot, n, counts, d, d2, d4),
              O.measure_bk(nb, N, 1.0, 1.0, True, kf, itot, n, counts, d, d2, d4))
    # cross bispectrum of two catalogues (bispectrumI4.f90:518-761)
    pos_b = _second_tracer(pos, 10)
    db = O.compute_density(pos_b.shape[1], O.assign_double_grid(N, 4, pos_b.shape[1], pos_b), N, 4,
                           True, 1000.0)
    itc = O.itot_count(nb, 1.0, 1.0, True, True)
    cc = B.count_triangles(nb, N, 1.0, 1.0, True, True, 4, itc)
    assert np.array_equal(cc, O.count_triangles_exact(nb, N, 1.0, 1.0, True, cross=True) * float(N) ** 3)
    _bk_cross_close(B.measure_bk_cross(nb, N, 1.0, 1.0, True, kf, 4, itc, n, cc, d, db),
                    O.measure_bk_cross(nb, N, 1.0, 1.0, True, kf, itc, n, cc, d, db), d, db, N, nb, kf)


def _bk_cross_close(mine, ref, da, db, N, nb, kf, rtol=1e-4):
    """k's and N_tr exact; B relative to max(|B|, sqrt(|P_a(j) P_a(l) P_b(i)| / N_tr))."""
    assert mine.shape == ref.shape == (5, ref.shape[1])
    for r in (0, 1, 2, 4):
        assert np.array_equal(mine[r], ref[r]), r
    kf3 = O.kf3_bk(kf)
    pw = []
    for dcl in (da, db):
        maps, nmodes = O.shell_maps(dcl, N, nb, 1.0, 1.0)
        pw.append(kf3 * np.sum(maps * maps, axis=1) / float(N) ** 3 / nmodes)
    l, j, i = (ref[0] - 1).astype(int), (ref[1] - 1).astype(int), (ref[2] - 1).astype(int)
    gauss = np.sqrt(np.abs(pw[0][j] * pw[0][l] * pw[1][i]) / np.maximum(ref[4], 1))
    # both are float32 records: allow their storage rounding on top of rtol
    store = _ulp32(ref[3]) + _ulp32(mine[3])
    err = np.maximum(np.abs(mine[3] - ref[3]) - store, 0) / np.maximum(np.abs(ref[3]), gauss)
    assert err.max() <= rtol, (err.max(), np.argmax(err))


# ------------------------------------------------------------------ reference-layout views


@pytest.mark.parametrize("inter", [True, False])
def test_device_objects_convert_to_reference_arrays(inter):
    import pbi4_amd as pb
    N, n = 16, 20000
    pos = (_clustered(n, 6)[0].T % 1000.0 / 1000.0).astype(np.float32)
    params = {"grid": N, "interpolation": 3, "interlacing": inter, "box": 1000.0,
              "kf": 2 * np.pi / 1000, "kN": np.pi / 1000 * N, "multipoles": True,
              "los": np.array([0.0, 0.0, 1.0])}
    den = pb.DENSITIES(params)
    npart, g = den.assign_grid(pos)
    ga = np.asarray(g)
    p64 = pos.astype(np.float64)
    gref = O.assign_double_grid(N, 3, n, p64) if inter else O.assign_single_grid(N, 3, n, p64)
    assert np.abs(ga - gref).max() <= 2e-6 * np.abs(gref).max()
    d = den.compute_density(npart, g)
    dref = O.compute_density(n, gref if inter else gref.astype(np.complex128), N, 3, inter, 1000.0)
    assert np.abs(np.asarray(d) - dref).max() <= 2e-5 * np.abs(dref).max()
    if inter:
        d2, d4 = den.compute_density_multipoles(d)
        r2, r4 = O.compute_density_multipoles(dref, N, 1000.0, [0.0, 0.0, 1.0])
        assert np.abs(np.asarray(d2) - r2).max() <= 2e-5 * np.abs(dref).max()
        assert np.abs(np.asarray(d4) - r4).max() <= 2e-5 * np.abs(dref).max()


# ------------------------------------------------------------------ fused entry and serving


def test_fused_pk_bk_equals_separate_calls():
    from pbi4_amd.core import Context
    ctx = Context(0)
    N, n, nb = 32, 50000, 8
    pos = torch.from_numpy(np.ascontiguousarray(_clustered(n, 21)[0].T % 1000.0)).cuda()
    g = ctx.assign(pos, N, 4, True, box=1000.0, normalised=False)
    d = ctx.density(g, N, 4, True, 1000.0, n)
    itot = O.itot_count(nb, 1.0, 1.0, True)
    ntr = ctx.count_triangles(nb, N, 1.0, 1.0, True, False, itot)
    counts = torch.from_numpy((ntr * float(N) ** 3).astype(np.float32)).cuda()
    for mult in (False, True):
        bd = ctx.bk_desc(nb, N, 1.0, 1.0, True, np.float32(2 * np.pi / 1000), itot, n, mult)
        pd = ctx.pk_desc(N, 1000.0, 1.0, 1.0, 16, mult)
        p1 = ctx.pk(d, N, 1000.0, 1.0, 1.0, 16, mult)
        b1 = ctx.bk(d, counts, bd)
        p2, b2 = ctx.pk_bk(d, counts, pd, bd)
        assert torch.equal(p1, p2) and torch.equal(b1, b2)


def test_serving_alone_and_in_the_mix():
    """Each catalogue measured alone (one slot) and inside a mixed continuous batch (4 slots,
    requests of different grids, orders and sizes arriving while others run) gives bitwise the
    same result: every kernel reduces in a fixed order."""
    from pbi4_amd.catalogue import zeldovich
    from pbi4_amd.core import Context
    from pbi4_amd.pipeline import Field, Measurement
    from pbi4_amd.serving import Request, ServingEngine
    ctx = Context(0)
    shapes = [Measurement(grid=32, order=4, kbin=2.0, kcenter=2.0, nbins_bisp=5),
              Measurement(grid=48, order=3, kbin=2.0, kcenter=2.0, nbins_bisp=6),
              Measurement(grid=64, order=2, kbin=2.0, kcenter=2.0, nbins_bisp=8,
                          fields=(Field(True, (0.0, 1.0, 0.0)),))]
    cats = [zeldovich(n, seed=50 + n, device="cuda") for n in (20, 28, 36, 24, 32)]
    jobs = [(shapes[i % 3], cats[i]) for i in range(5)]
    alone = []
    eng1 = ServingEngine(ctx, slots=1)
    for i, (m, (p, v)) in enumerate(jobs):
        alone.append(eng1.run([Request(i, m, p, v)])[0].result)
    eng4 = ServingEngine(ctx, slots=4)
    reqs = [Request(i, m, p, v, arrival=0.002 * i) for i, (m, (p, v)) in enumerate(jobs)]
    done = {r.rid: r.result for r in eng4.run(reqs)}
    for i in range(5):
        for f in alone[i]:
            assert np.array_equal(alone[i][f][0], done[i][f][0])
            assert np.array_equal(alone[i][f][1], done[i][f][1])
    # and one of them against the FP64 oracle
    m, (p, v) = jobs[0]
    res = done[0][Field(False)]
    N = m.grid
    # the reference's normalisation in float32 (pbi4/pbi4.py:241-243), as the device applies it
    box = np.float32(1000.0)
    pos = np.minimum(((p.cpu().numpy() + box) % box) / box, np.float32(1.0))
    npart = pos.shape[1]
    d = O.compute_density(npart, O.assign_double_grid(N, 4, npart, pos.astype(np.float64)), N, 4, True, 1000.0)
    nb, itot = m.nb_bisp(), m.itot()
    rc = (O.count_triangles_exact(nb, N, 2.0, 2.0, True) * float(N) ** 3).astype(np.float32)
    bref = O.measure_bk(nb, N, 2.0, 2.0, True, m.kf(), itot, npart, rc, d)
    # 3 R >= N here: the maps are on the full grid, where split-BF16 sums are held to 3e-4
    # (the full-grid bound of tests/test_gpu_configs.py); FP32 sums to 1e-4
    assert ctx.maps_grid(ctx.bk_desc(nb, N, 2.0, 2.0, True, m.kf(), itot, npart)) == N
    _bk_close(res[1].T, bref, rtol=3e-4)
