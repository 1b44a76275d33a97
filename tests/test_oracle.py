"""Pins the FP64 oracle before anything is compared to it.

Each check is chosen so a plausible restatement mistake fails it:
  * reference-held data: ps.dat Nmodes, bs.dat triangle list and N_tr (tests/golden);
  * the reference itself: powerI4.f90 compiled unchanged into oracle/_ref (skipped if absent);
  * closed forms: B-spline weights, mass conservation, L_l(mu), a three-plane-wave bispectrum,
    Parseval for pow(s), brute-force lattice counts.
"""
import ctypes
import os

import numpy as np
import pytest

from oracle import pbi4_oracle as O

REF_SO = os.path.join(os.path.dirname(O.__file__), "_ref", "libpowerI4_ref.so")


def _ref():
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    return ctypes.CDLL(REF_SO)


def _load_ps(golden_dir):
    a = np.loadtxt(os.path.join(golden_dir, "ps_N64_kb1_kc1.txt"))
    return a[:, 0], a[:, 1].astype(np.int64), a[:, 2]


def _load_bs(golden_dir):
    return np.loadtxt(os.path.join(golden_dir, "bs_N64_kb1_kc1.txt")).astype(np.int64)


def _rand_pos(n, seed, clustered=False):
    rng = np.random.default_rng(seed)
    if not clustered:
        return rng.random((3, n)).astype(np.float32)
    c = rng.random((3, 8))
    k = rng.integers(0, 8, n)
    p = (c[:, k] + 0.05 * rng.standard_normal((3, n))) % 1.0
    return np.minimum(p.astype(np.float32), np.float32(1.0) - np.float32(2 ** -24))


# ------------------------------------------------------------------ derived parameters


def test_derived_params_known_answers():
    base = dict(box=1000.0, kbin=1.0, kcenter=1.0, iopen=True, om=0.3, ol=0.7, z=0.5,
                los=[0.0, 0.0, 1.0], cross=False)
    d = O.derived_params(dict(base, grid=128))         # Tutorials/Notebook/config.json:4-33
    assert (d["nbins_ps"], d["nbins_bisp"], d["itot"]) == (64, 42, 7714)
    d = O.derived_params(dict(base, grid=64))          # bs.dat
    assert (d["nbins_ps"], d["nbins_bisp"], d["itot"]) == (32, 21, 1156)
    assert abs(d["Hz"] - 130.86252328302402) < 1e-9
    d = O.derived_params(dict(base, grid=64, kbin=2.0, kcenter=2.0))  # Script/config.json itot
    assert d["itot"] == 1156 or d["nbins_bisp"] == 10
    assert O.itot_count(16, 2.0, 2.0, True) == 564
    assert O.itot_count(48, 1.0, 1.0, True) == 11228
    assert O.itot_count(32, 2.0, 2.0, True) == 3624
    assert O.itot_count(8, 2.0, 2.0, True) == 98


def test_golden_kvec_and_psn(golden_dir):
    k, _, psn = _load_ps(golden_dir)
    d = O.derived_params(dict(box=1000.0, grid=64, kbin=1.0, kcenter=1.0, iopen=True, om=0.3,
                              ol=0.7, z=0.5, los=[0, 0, 1], cross=False))
    assert np.array_equal(d["kvec"].astype(np.float32), k.astype(np.float32))
    npart = 356467                                     # Sancho catalogue size (notebook output)
    p = np.float32(1.0 / (npart * (2.0 * np.pi / 1000.0) ** 3))
    assert np.all(psn.astype(np.float32) == p)


# ------------------------------------------------------------------ mode counts / triangles


def test_golden_pk_nmodes(golden_dir):
    _, nmodes, _ = _load_ps(golden_dir)
    imk = O.pk_bin_index(64, 1000.0, 1.0, 1.0)
    co = np.bincount(imk.ravel(), minlength=34)[1:33]
    assert np.array_equal(co, nmodes)


def test_golden_triangle_list(golden_dir):
    bs = _load_bs(golden_dir)
    tl = O.triangle_list(21, 1.0, 1.0, True)
    tk = 1.0 + (tl - 1) * 1.0
    assert np.array_equal(tk.astype(np.int64), bs[:, :3])


def test_golden_ntr_exact_integer(golden_dir):
    bs = _load_bs(golden_dir)
    ntr = O.count_triangles_exact(21, 64, 1.0, 1.0, True)
    assert np.array_equal(ntr, bs[:, 3])


@pytest.mark.slow
def test_golden_ntr_fft_form(golden_dir):
    bs = _load_bs(golden_dir)
    c = O.count_triangles(21, 64, 1.0, 1.0, True)
    ntr = c / 64.0 ** 3
    assert np.max(np.abs(ntr - np.rint(ntr))) < 1e-6
    assert np.array_equal(np.rint(ntr).astype(np.int64), bs[:, 3])
    # the value measure_bk divides by: REAL(4) counts, bout(9) = real(counts/N^3)
    assert np.array_equal((c.astype(np.float32) / np.float32(64 ** 3)), bs[:, 3].astype(np.float32))


def test_exact_counts_match_fft_counts_with_aliasing():
    # N=12, 4 shells of kbin=1 reach |q|=4.5 > N/3: the modular (aliased) closure counts
    for (N, nb, kb, kc, io) in [(12, 4, 1.0, 1.0, True), (16, 4, 2.0, 2.0, True),
                                (12, 3, 1.5, 1.0, False)]:
        e = O.count_triangles_exact(nb, N, kb, kc, io)
        f = O.count_triangles(nb, N, kb, kc, io) / N ** 3
        assert np.array_equal(e, np.rint(f).astype(np.int64)), (N, nb)


def test_exact_counts_brute_force_tiny():
    # pure-python enumeration of all (q1, q2, q3) triples on an 8^3 grid
    N, nb, kb, kc = 8, 2, 1.0, 1.0
    idist = O.shell_index_grid(N, kb, kc)
    pts = {s: [tuple(p) for p in np.argwhere(idist == s)] for s in (1, 2)}
    tl = O.triangle_list(nb, kb, kc, True)
    e = O.count_triangles_exact(nb, N, kb, kc, True)
    for m, (l, j, i) in enumerate(tl):
        cnt = 0
        for a in pts[i]:
            for b in pts[j]:
                c = tuple((-(a[d] + b[d])) % N for d in range(3))
                cnt += idist[c] == l
        assert cnt == e[m]


# ------------------------------------------------------------------ assignment


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_assign_against_reference(order):
    ref = _ref()
    N, n = 16, 3000
    pos = _rand_pos(n, 10 + order).astype(np.float64)
    w = np.ones(n)
    posF = np.asfortranarray(pos)
    for inter in (0, 1):
        if inter:
            out = np.zeros((N, N, N), dtype=np.complex128, order="F")
            ref.ref_assign_double_grid(N, order, n, posF.ctypes.data_as(ctypes.c_void_p),
                                       w.ctypes.data_as(ctypes.c_void_p),
                                       out.ctypes.data_as(ctypes.c_void_p))
            mine = O.assign_double_grid(N, order, n, pos, w)
        else:
            out = np.zeros((N, N, N), dtype=np.float64, order="F")
            ref.ref_assign_single_grid(N, order, n, posF.ctypes.data_as(ctypes.c_void_p),
                                       w.ctypes.data_as(ctypes.c_void_p),
                                       out.ctypes.data_as(ctypes.c_void_p))
            mine = O.assign_single_grid(N, order, n, pos, w)
        assert np.allclose(mine, out, rtol=1e-13, atol=1e-13), (order, inter)


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_assign_threads_bit_identical_to_serial(order):
    # the z-slab split keeps every cell's additions in catalogue order: any thread count gives
    # the serial loop's result bit for bit (odd N, more threads than planes, edge positions)
    n = 40000
    pos = _rand_pos(n, 30 + order).astype(np.float64)
    pos[:, :4] = np.array([0.0, 1.0, 0.5, 1.0 - 2.0 ** -40])[None, :]
    for N in (16, 7, 3):
        for fn in (O.assign_double_grid, O.assign_single_grid):
            serial = fn(N, order, n, pos, workers=1)
            for t in (2, 5, 8):
                assert np.array_equal(fn(N, order, n, pos, workers=t), serial), (N, t)


@pytest.mark.parametrize("order", [1, 3, 4])
def test_assign_single_particle_closed_form(order):
    N = 16
    rng = np.random.default_rng(order)
    for _ in range(20):
        p = rng.random(3)
        g = O.assign_double_grid(N, order, 1, p.reshape(3, 1))
        u = N * p
        for grid_sel, shift in ((0, 0.0), (1, -0.5)):       # grid 2 samples u - 1/2
            val = g.real if grid_sel == 0 else g.imag
            exp = np.ones((N, N, N))
            for d in range(3):
                c = np.arange(N)
                dist = (u[d] + shift - c + N / 2) % N - N / 2
                w1 = np.array([O.bspline_weights_closed_form(order, x) for x in dist])
                shape = [1, 1, 1]
                shape[d] = N
                exp = exp * w1.reshape(shape)
            if order == 1:   # NGP ties are measure-zero; compare mass location only
                assert np.isclose(val.sum(), 1.0)
                assert np.argmax(val) == np.argmax(exp)
            else:
                assert np.allclose(val, exp, atol=1e-12)


def test_assign_cic_double_grid_is_shifted_one_cell():
    # reference CIC double grid deposits at floor(u)+1, floor(u)+2 (powerI4.f90:208-224)
    N = 16
    p = np.array([[0.3], [0.55], [0.9]])
    g = O.assign_double_grid(N, 2, 1, p).real
    u = N * p[:, 0]
    base = np.floor(u).astype(int) + 1
    h = u - np.floor(u)
    assert np.isclose(g[base[0] % N, base[1] % N, base[2] % N], np.prod(1 - h))
    assert np.isclose(g.sum(), 1.0)


def test_assign_mass_conservation():
    pos = _rand_pos(5000, 3, clustered=True)
    for order in (2, 3, 4):
        g = O.assign_double_grid(32, order, 5000, pos)
        assert np.isclose(g.real.sum(), 5000.0, rtol=1e-12)
        assert np.isclose(g.imag.sum(), 5000.0, rtol=1e-12)


# ------------------------------------------------------------------ fcomb / window / multipoles


def test_window_table_closed_form():
    N = 32
    for p in (1, 2, 3, 4):
        w = O.window_table(N, p)
        k = np.arange(N // 2 + 1) * np.float64(O.TWOPI) / N / 2
        exp = np.ones_like(k)
        exp[1:] = (np.sin(k[1:]) / k[1:]) ** p
        assert np.allclose(w, exp, rtol=1e-14)


@pytest.mark.parametrize("inter", [0, 1])
def test_fcomb_against_reference(inter):
    ref = _ref()
    N = 16
    rng = np.random.default_rng(5)
    fd = np.asfortranarray(rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)))
    out = np.zeros((N, N, N), dtype=np.complex128, order="F")
    ref.ref_fcomb(N, 3, inter, ctypes.c_float(1000.0), fd.ctypes.data_as(ctypes.c_void_p),
                  out.ctypes.data_as(ctypes.c_void_p))
    mine = O.fcomb(N, 3, inter, 1000.0, fd)
    assert np.allclose(mine, out, rtol=1e-13, atol=1e-9 * np.abs(out).max())


@pytest.mark.parametrize("L,G", [(16, 8), (12, 12), (16, 10), (9, 5)])
def test_reduce_array_against_reference(L, G):
    ref = _ref()
    rng = np.random.default_rng(L * 31 + G)
    din = np.asfortranarray(rng.standard_normal((L, L, L)) + 1j * rng.standard_normal((L, L, L)))
    out = np.zeros((G, G, G), dtype=np.complex128, order="F")
    ref.ref_reduce_array(L, din.ctypes.data_as(ctypes.c_void_p), G, out.ctypes.data_as(ctypes.c_void_p))
    assert np.array_equal(O.reduce_array(L, din, G), out)


def test_reduce_array_identity_and_plane_waves():
    # N = L keeps the x <= N/2 half unchanged; a field of plane waves with |k_i| < N/2 sampled on
    # L^3 reduces to the same waves' DFT on N^3 (times (L/N)^3, the DFT sum length)
    L, N = 24, 12
    rng = np.random.default_rng(4)
    a = rng.standard_normal((L, L, L)) + 1j * rng.standard_normal((L, L, L))
    d = O.reduce_array(L, a, L)
    assert np.array_equal(d[:L // 2 + 1], a[:L // 2 + 1]) and not d[L // 2 + 1:].any()

    def field(n):
        x = np.arange(n) / n
        X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
        f = np.zeros((n, n, n))
        for (kx, ky, kz), amp in (((1, 2, -3), 1.0), ((5, -5, 0), 0.5), ((2, 0, 5), -2.0)):
            f += amp * np.cos(2 * np.pi * (kx * X + ky * Y + kz * Z) + 0.3 * kx)
        return np.fft.fftn(f)

    red = O.reduce_array(L, field(L), N) * (N / L) ** 3
    exp = field(N)
    h = N // 2 + 1
    assert np.allclose(red[:h], exp[:h], atol=1e-9 * np.abs(exp).max())


def test_reduce_array_nyquist_takes_the_negative_frequency():
    # the reference's store order: on the Nyquist index of an even N the source at L - N/2
    # (frequency -N/2) overwrites the one at N/2
    L, N = 16, 8
    a = np.zeros((L, L, L), dtype=np.complex128)
    a[1, N // 2, 2] = 1.0
    a[1, L - N // 2, 2] = 2.0
    a[1, 3, N // 2] = 3.0
    a[1, 3, L - N // 2] = 4.0
    a[2, N // 2, N // 2] = 5.0
    a[2, L - N // 2, L - N // 2] = 6.0
    d = O.reduce_array(L, a, N)
    assert d[1, N // 2, 2] == 2.0 and d[1, 3, N // 2] == 4.0 and d[2, N // 2, N // 2] == 6.0


def test_fcomb_interlaced_closed_form_interior():
    # for real grids g1, g2: delta = (G1 + phi G2) / (2 kF3 W) away from self-conjugate planes
    N = 16
    rng = np.random.default_rng(9)
    g1, g2 = rng.random((N, N, N)), rng.random((N, N, N))
    F = np.fft.fftn(g1 + 1j * g2)
    D = O.fcomb(N, 4, 1, 1000.0, F)
    G1, G2 = np.fft.fftn(g1), np.fft.fftn(g2)
    W = O.window_table(N, 4)
    kF3 = O.kf3_f32(1000.0)
    pi_ = np.float64(O.TWOPI) / 2
    for (a, b, c) in [(1, 2, 3), (3, 14, 5), (7, 1, 9), (2, 10, 13)]:
        sy = b if b <= N // 2 else b - N
        sz = c if c <= N // 2 else c - N
        phi = np.exp(-1j * pi_ * (a + sy + sz) / N)
        exp = (G1[a, b, c] + phi * G2[a, b, c]) / (2 * kF3 * W[a] * W[abs(sy)] * W[abs(sz)])
        assert np.isclose(D[a, b, c], exp, rtol=1e-10)


def test_q2k_q4k_against_reference():
    ref = _ref()
    N = 8
    rng = np.random.default_rng(2)
    dk = np.asfortranarray(rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)))
    for (i, j) in [(1, 1), (3, 2), (3, 3)]:
        out = np.zeros((N, N, N), dtype=np.complex128, order="F")
        ref.ref_compute_q2k(N, ctypes.c_float(1000.0), dk.ctypes.data_as(ctypes.c_void_p), i, j,
                            out.ctypes.data_as(ctypes.c_void_p))
        assert np.allclose(O.compute_q2k(N, 1000.0, dk, i, j), out, rtol=1e-14, atol=1e-14)
    out = np.zeros((N, N, N), dtype=np.complex128, order="F")
    ref.ref_compute_q4k(N, ctypes.c_float(1000.0), dk.ctypes.data_as(ctypes.c_void_p),
                        3, 2, 2, 1, out.ctypes.data_as(ctypes.c_void_p))
    assert np.allclose(O.compute_q4k(N, 1000.0, dk, 3, 2, 2, 1), out, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("los", [[0, 0, 1.0], [0.0, 2.0, 1.0]])
def test_multipoles_literal_equals_legendre(los):
    N = 16
    pos = _rand_pos(4000, 4, clustered=True)
    dcl = O.compute_density(4000, O.assign_double_grid(N, 3, 4000, pos), N, 3, 1, 1000.0)
    l = np.array(los) / (np.array(los) @ np.array(los))          # pbi4.py:103
    d2, d4 = O.compute_density_multipoles(dcl, N, 1000.0, l)
    L2, L4 = O.legendre_factors(N, l)
    h = N // 2 + 1
    scale = np.abs(dcl[:h]).max()
    assert np.allclose(d2[:h], L2 * dcl[:h], atol=1e-12 * scale)
    assert np.allclose(d4[:h], L4 * dcl[:h], atol=1e-12 * scale)


# ------------------------------------------------------------------ P(k)


def test_measure_pk_against_reference():
    ref = _ref()
    N = 16
    pos = _rand_pos(4000, 6, clustered=True)
    dcl = np.asfortranarray(O.compute_density(4000, O.assign_double_grid(N, 4, 4000, pos),
                                              N, 4, 1, 1000.0))
    d2, d4 = (np.asfortranarray(x) for x in O.compute_density_multipoles(dcl, N, 1000.0, [0, 0, 1.0]))
    nb = 8
    for kb, kc in [(1.0, 1.0), (2.0, 2.0), (1.5, 0.5)]:
        out = np.zeros((5, nb), order="F")
        args = (N, ctypes.c_float(1000.0), ctypes.c_float(kb), ctypes.c_float(kc), nb)
        ref.ref_measure_pk(*args, dcl.ctypes.data_as(ctypes.c_void_p),
                           dcl.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p))
        mine = O.measure_pk(N, 1000.0, kb, kc, nb, dcl, dcl)
        assert np.array_equal(mine[2], out[2])                  # mode counts bit-exact
        assert np.allclose(mine, out, rtol=1e-13, equal_nan=True)
        out = np.zeros((5, nb), order="F")
        ref.ref_measure_pk_multipoles(*args, dcl.ctypes.data_as(ctypes.c_void_p),
                                      d2.ctypes.data_as(ctypes.c_void_p),
                                      d4.ctypes.data_as(ctypes.c_void_p),
                                      out.ctypes.data_as(ctypes.c_void_p))
        mine = O.measure_pk_multipoles(N, 1000.0, kb, kc, nb, dcl, d2, d4)
        assert np.array_equal(mine[4], out[4])
        assert np.allclose(mine, out, rtol=1e-13, equal_nan=True)


def test_pk_bin_index_matches_measure_pk_counts():
    N = 16
    for kb, kc in [(1.0, 1.0), (2.0, 2.0), (1.5, 0.5), (2.0, 1.0)]:
        imk = O.pk_bin_index(N, 1000.0, kb, kc)
        d = np.ones((N, N, N), dtype=np.complex128)
        co = O.measure_pk(N, 1000.0, kb, kc, 10, d, d)[2]
        exp = np.bincount(imk.ravel(), minlength=12)[1:11].astype(np.float64)
        assert np.array_equal(co, exp)
        # the blocked form (planes of x) used at large N gives the same counts
        for planes in (1, 3, 32):
            assert np.array_equal(O.pk_mode_counts(N, 1000.0, kb, kc, 10, planes), exp)


def test_pk_mode_counts_golden(golden_dir):
    _, nmodes, _ = _load_ps(golden_dir)
    assert np.array_equal(O.pk_mode_counts(64, 1000.0, 1.0, 1.0, len(nmodes), planes=5), nmodes)


# ------------------------------------------------------------------ bispectrum


def _plane_wave_field(N, qs, amps):
    """dcl with delta at +-q for each q (Hermitian), stored over the full grid."""
    D = np.zeros((N, N, N), dtype=np.complex128)
    for q, a in zip(qs, amps):
        D[q[0] % N, q[1] % N, q[2] % N] += a
        D[(-q[0]) % N, (-q[1]) % N, (-q[2]) % N] += np.conj(a)
    return D


def test_bispectrum_three_plane_waves_closed_form():
    N, nb, kb, kc = 16, 5, 1.0, 1.0
    q1, q2 = np.array([1, 2, 0]), np.array([-3, 0, 1])      # |q1|^2=5 (shell 2), |q2|^2=10 (3)
    q3 = -(q1 + q2)                                          # (2,-2,-1): |q3|^2 = 9 (shell 3)
    a = [0.7 + 0.2j, -0.3 + 0.9j, 1.1 - 0.4j]
    dcl = _plane_wave_field(N, [q1, q2, q3], a)
    tl = O.triangle_list(nb, kb, kc, True)
    counts = O.count_triangles_exact(nb, N, kb, kc, True).astype(np.float64) * N ** 3
    kf = np.float32(2 * np.pi / 1000.0)
    bout = O.measure_bk(nb, N, kb, kc, True, kf, len(tl), 10 ** 18, counts, dcl)
    kf3 = O.kf3_bk(kf)
    raw = bout[6] - bout[7]
    sh = {2: [0], 3: [1, 2]}
    for m, (l, j, i) in enumerate(tl):
        # enumerate ordered (q_i, q_j, q_l) from {+-q1,+-q2,+-q3} in shells (i, j, l) summing to 0
        vecs = [(q1, a[0]), (q2, a[1]), (q3, a[2])]
        allv = vecs + [(-v, np.conj(c)) for v, c in vecs]
        s = 0.0
        for va, ca in allv:
            for vb, cb in allv:
                for vc, cc in allv:
                    if np.any(va + vb + vc):
                        continue
                    sa = O.shell_of_d2(va @ va, kb, kc)
                    sb = O.shell_of_d2(vb @ vb, kb, kc)
                    sc = O.shell_of_d2(vc @ vc, kb, kc)
                    if (sa, sb, sc) == (i, j, l):
                        s += (ca * cb * cc).real
        exp = kf3 * s * N ** 3 / counts[m]
        assert np.isclose(raw[m], exp, rtol=1e-10, atol=1e-12 * abs(kf3 * N ** 3 / counts[m])), \
            (l, j, i)
    del sh


def test_bk_cross_of_one_field_equals_auto_raw():
    # sum_x M_i M_j M_l is symmetric: the cross estimate of a field with itself equals the auto
    # estimate without shot noise (B0 + BSN - BSN) of the sorted triangle, and every triangle of
    # the cross loop sorts into the auto loop; the exact counts agree the same way
    N, npart, nb = 16, 5000, 5
    pos = _rand_pos(npart, 21, clustered=True)
    dcl = O.compute_density(npart, O.assign_double_grid(N, 3, npart, pos), N, 3, 1, 1000.0)
    kf = np.float32(2 * np.pi / 1000.0)
    ta = O.triangle_list(nb, 1.0, 1.0, True)
    tc = O.triangle_list(nb, 1.0, 1.0, True, cross=True)
    assert len(tc) == O.itot_count(nb, 1.0, 1.0, True, True) > len(ta)
    ca = O.count_triangles_exact(nb, N, 1.0, 1.0, True).astype(np.float64) * N ** 3
    cc = O.count_triangles_exact(nb, N, 1.0, 1.0, True, cross=True).astype(np.float64) * N ** 3
    auto = O.measure_bk(nb, N, 1.0, 1.0, True, kf, len(ta), npart, ca, dcl)
    cross = O.measure_bk_cross(nb, N, 1.0, 1.0, True, kf, len(tc), npart, cc, dcl, dcl)
    where = {tuple(t): m for m, t in enumerate(ta)}
    for m, (l, j, i) in enumerate(tc):
        k = where[tuple(sorted((l, j, i), reverse=True))]
        assert cc[m] == ca[k]
        assert np.isclose(cross[3, m], auto[6, k] - auto[7, k], rtol=1e-10, atol=1e-12 * abs(auto[7, k]))
        assert (cross[0, m], cross[1, m], cross[2, m]) == (float(l), float(j), float(i))


def test_bk_cross_plane_waves_closed_form():
    # field A holds waves at q1, q2, field B at q3 = -(q1 + q2): only B in shell i and A in
    # shells j, l contribute, so a swapped role of the two fields fails this
    N, nb, kb, kc = 16, 5, 1.0, 1.0
    q1, q2 = np.array([1, 2, 0]), np.array([-3, 0, 1])      # shells 2, 3
    q3 = -(q1 + q2)                                          # shell 3
    a = [0.7 + 0.2j, -0.3 + 0.9j, 1.1 - 0.4j]
    fa = _plane_wave_field(N, [q1, q2], a[:2])
    fb = _plane_wave_field(N, [q3], a[2:])
    tc = O.triangle_list(nb, kb, kc, True, cross=True)
    cc = O.count_triangles_exact(nb, N, kb, kc, True, cross=True).astype(np.float64) * N ** 3
    kf = np.float32(2 * np.pi / 1000.0)
    bout = O.measure_bk_cross(nb, N, kb, kc, True, kf, len(tc), 1, cc, fa, fb)
    kf3 = O.kf3_bk(kf)
    va = [(q1, a[0]), (q2, a[1])]
    va = va + [(-v, np.conj(c)) for v, c in va]
    vb = [(q3, a[2]), (-q3, np.conj(a[2]))]
    for m, (l, j, i) in enumerate(tc):
        s = 0.0
        for vi, ci in vb:
            for vj, cj in va:
                for vl, cl in va:
                    if np.any(vi + vj + vl):
                        continue
                    shells = tuple(O.shell_of_d2(v @ v, kb, kc) for v in (vi, vj, vl))
                    if shells == (i, j, l):
                        s += (ci * cj * cl).real
        exp = kf3 * s * N ** 3 / cc[m]
        assert np.isclose(bout[3, m], exp, rtol=1e-10, atol=1e-12 * abs(kf3 * N ** 3 / cc[m])), (l, j, i)
    assert np.any(np.abs(bout[3]) > 0)


def test_pow_parseval_and_pk_invariant():
    # pow(i) from the k maps equals half-space P0 - PSN on the same shells (kb = kc = 1 has no
    # bin-edge ties), up to the REAL(4) twopi vs 2*pi/box constants (~2e-6)
    N, npart = 32, 20000
    pos = _rand_pos(npart, 8, clustered=True)
    dcl = O.compute_density(npart, O.assign_double_grid(N, 3, npart, pos), N, 3, 1, 1000.0)
    nb = 8
    tl = O.triangle_list(nb, 1.0, 1.0, True)
    counts = O.count_triangles_exact(nb, N, 1.0, 1.0, True).astype(np.float64) * N ** 3
    kf = np.float32(2 * np.pi / 1000.0)
    bout = O.measure_bk(nb, N, 1.0, 1.0, True, kf, len(tl), npart, counts, dcl)
    pk = O.measure_pk(N, 1000.0, 1.0, 1.0, nb, dcl, dcl)
    psn = 1.0 / (npart * (2 * np.pi / 1000.0) ** 3)
    pw = np.zeros(nb)
    for m, (l, j, i) in enumerate(tl):
        pw[l - 1], pw[j - 1], pw[i - 1] = bout[3, m], bout[4, m], bout[5, m]
    assert np.allclose(pw, pk[1] - psn, rtol=5e-6)


def test_bk_multipoles_plane_waves_closed_form():
    # delta holds waves at q1 (shell 2) and q2 (shell 3) only; delta_2 and delta_4 hold waves of
    # different amplitudes at q3 = -(q1 + q2) (shell 4) only.  The reference attaches the L_2 / L_4
    # map to the largest bin l and scales by 5 and 9 (bispectrumI4.f90:484-488, 499-500), so B2 and
    # B4 are nonzero on the triangle (4, 3, 2) alone and equal the brute-force mode sums there.
    # The ell map on bin i or j, a wrong factor, or delta_2 and delta_4 swapped all fail this.
    N, nb, kb, kc = 16, 5, 1.0, 1.0
    q1, q2 = np.array([2, 0, 0]), np.array([0, 3, 0])
    q3 = -(q1 + q2)
    assert [O.shell_of_d2(q @ q, kb, kc) for q in (q1, q2, q3)] == [2, 3, 4]
    a = [0.7 + 0.2j, -0.3 + 0.9j]
    b = {2: 1.1 - 0.4j, 4: -0.6 + 0.5j}
    dcl = _plane_wave_field(N, [q1, q2], a)
    d2 = _plane_wave_field(N, [q3], [b[2]])
    d4 = _plane_wave_field(N, [q3], [b[4]])
    tl = O.triangle_list(nb, kb, kc, True)
    counts = O.count_triangles_exact(nb, N, kb, kc, True).astype(np.float64) * N ** 3
    kf = np.float32(2 * np.pi / 1000.0)
    bout = O.measure_bk(nb, N, kb, kc, True, kf, len(tl), 10 ** 18, counts, dcl, d2, d4)
    kf3 = O.kf3_bk(kf)
    va = [(q1, a[0]), (q2, a[1])]
    va = va + [(-v, np.conj(c)) for v, c in va]
    hits = 0
    for row, ell, fac in ((9, 2, 5.0), (10, 4, 9.0)):
        vl = [(q3, b[ell]), (-q3, np.conj(b[ell]))]
        for m, (l, j, i) in enumerate(tl):
            s = 0.0
            for vi, ci in va:
                for vj, cj in va:
                    for vk, ck in vl:
                        if np.any(vi + vj + vk):
                            continue
                        shells = tuple(O.shell_of_d2(v @ v, kb, kc) for v in (vi, vj, vk))
                        if shells == (i, j, l):
                            s += (ci * cj * ck).real
            exp = fac * kf3 * s * N ** 3 / counts[m]
            hits += s != 0.0
            assert np.isclose(bout[row, m], exp, rtol=1e-10,
                              atol=1e-12 * abs(kf3 * N ** 3 / counts[m])), (row, l, j, i)
    assert hits == 2                                         # (4, 3, 2) in B2 and in B4
    m = [tuple(t) for t in tl].index((4, 3, 2))
    # the two waves of delta pair once each way round with the q3 wave: 2 Re(a1 a2 b_ell)
    for row, ell, fac in ((9, 2, 5.0), (10, 4, 9.0)):
        exp = fac * kf3 * 2 * (a[0] * a[1] * b[ell]).real * N ** 3 / counts[m]
        assert np.isclose(bout[row, m], exp, rtol=1e-10)


@pytest.mark.parametrize("multipoles", [False, True])
def test_bk_subset_equals_full_columns(multipoles):
    # the subset form builds maps of the sampled triangles' shells only; every column must be
    # the full measure_bk column bit for bit (same maps, sums and row formulas)
    N, npart, nb = 16, 8000, 6
    pos = _rand_pos(npart, 41, clustered=True)
    dcl = O.compute_density(npart, O.assign_double_grid(N, 4, npart, pos), N, 4, 1, 1000.0)
    tl = O.triangle_list(nb, 1.0, 1.0, True)
    counts = O.count_triangles_exact(nb, N, 1.0, 1.0, True).astype(np.float64) * N ** 3
    kf = np.float32(2 * np.pi / 1000.0)
    extra = O.compute_density_multipoles(dcl, N, 1000.0, [0, 0, 1.0]) if multipoles else (None, None)
    full = O.measure_bk(nb, N, 1.0, 1.0, True, kf, len(tl), npart, counts, dcl, *extra)
    for rows_m in ([0], [len(tl) - 1, 3, 17], list(range(0, len(tl), 5))):
        sub = O.measure_bk_subset(nb, N, 1.0, 1.0, True, kf, len(tl), npart, counts, dcl, rows_m,
                                  *extra)
        assert np.array_equal(sub, full[:, rows_m]), rows_m


def test_bk_multipoles_real_space_limit():
    # with los along z, B2 and B4 use L2/L4-weighted maps; the l=0 rows must equal measure_bk
    N, npart, nb = 16, 6000, 4
    pos = _rand_pos(npart, 12, clustered=True)
    dcl = O.compute_density(npart, O.assign_double_grid(N, 4, npart, pos), N, 4, 1, 1000.0)
    d2, d4 = O.compute_density_multipoles(dcl, N, 1000.0, [0, 0, 1.0])
    tl = O.triangle_list(nb, 1.0, 1.0, True)
    counts = O.count_triangles_exact(nb, N, 1.0, 1.0, True).astype(np.float64) * N ** 3
    kf = np.float32(2 * np.pi / 1000.0)
    b9 = O.measure_bk(nb, N, 1.0, 1.0, True, kf, len(tl), npart, counts, dcl)
    b11 = O.measure_bk(nb, N, 1.0, 1.0, True, kf, len(tl), npart, counts, dcl, d2, d4)
    assert np.array_equal(b9, b11[:9])
    assert np.all(np.isfinite(b11))
