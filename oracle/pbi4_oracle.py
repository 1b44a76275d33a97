"""FP64 CPU oracle for the PBI4 measurement path -- TEST INFRASTRUCTURE ONLY.

This module is the checker for the HIP implementation in ``pbi4_amd``.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import it; the
product package never does.  It restates the reference routines literally (FP64, the
reference's operation order, Fortran NINT/INT/MOD semantics, twopi = REAL(4) 6.28319):

  derived_params ........... pbi4/pbi4.py:86-119
  assign_single_grid ....... powerI4.f90:68-140      (C: oracle_kernels.c orc_assign)
  assign_double_grid ....... powerI4.f90:142-249     (C: oracle_kernels.c orc_assign)
  compute_density .......... pbi4/pbi4.py:304-341    (scipy.fft.fftn + fcomb)
  fcomb .................... powerI4.f90:251-378     (vectorised; the 8 ordered stores kept)
  reduce_array ............. powerI4.f90:380-407     (the 4 ordered stores kept)
  catalogue_positions ...... pbi4/pbi4.py:151-243    (the RSD shift, wrap and normalisation lines)
  compute_q2k / compute_q4k  powerI4.f90:409-508
  compute_density_multipoles pbi4/pbi4.py:343-388   (literal ifftn/fftn path)
  measure_pk ............... powerI4.f90:510-594     (C: orc_measure_pk mode 0)
  measure_pk_multipoles .... powerI4.f90:596-693     (C: orc_measure_pk mode 1)
  shell index / k maps ..... bispectrumI4.f90:65-157
  measure_bk ............... bispectrumI4.f90:1-198
  measure_bk_multipoles .... bispectrumI4.f90:200-516
  measure_bk_subset ........ the same, for a subset of triangles (maps of their shells only)
  count_triangles .......... bispectrumI4.f90:763-915 (FFT form) and an exact integer form

How each function is pinned (checked by tests/test_oracle.py):
  * assignment, fcomb, reduce_array, q2k/q4k, measure_pk(_multipoles): against the reference powerI4.f90
    itself, compiled into oracle/_ref by oracle/Makefile, plus closed forms (B-spline weights,
    mass conservation, window sinc^p).
  * shell index / mode counts / triangle enumeration / N_tr: against the reference's saved
    outputs Tutorials/testdata/ps.dat (Nmodes) and bs.dat (triangles, N_tr), committed as
    tests/golden/*.txt, and against brute-force lattice enumeration at small N.
  * k maps and pow(s): Parseval (pow(s) = half-space P0 - PSN on the same shells).
  * measure_bk(_multipoles): closed-form bispectrum of a three-plane-wave field, symmetry of
    the triple sum, and the literal FFT-based formula at small N.
  * compute_density_multipoles: against the closed form delta_l = L_l(mu) delta.
  * catalogue_positions: the reference's numpy lines run as written (numpy is the arithmetic
    being restated), plus the invariants 0 <= x <= 1 and x * box = position mod box.
"""
import ctypes
import os
import subprocess

import numpy as np
import scipy.fft

HERE = os.path.dirname(os.path.abspath(__file__))
TWOPI = np.float32(6.28319)          # REAL(4), parameter :: twopi = 6.28319
TINY = np.finfo(np.float64).tiny     # tiny(1.d0)

# --------------------------------------------------------------------------------------------
# helpers


def fnint(x):
    """Fortran NINT for REAL(8): round half away from zero, exactly (no x+0.5 double rounding)."""
    x = np.asarray(x, dtype=np.float64)
    t = np.trunc(x)
    f = x - t
    return (t + (f >= 0.5) - (f <= -0.5)).astype(np.int64)


def _lib():
    path = os.path.join(HERE, "liboracle.so")
    if not os.path.exists(path):
        subprocess.check_call(["make", "-s", "-f", os.path.join(HERE, "Makefile"),
                               os.path.join(HERE, "liboracle.so")])
    lib = ctypes.CDLL(path)
    d = ctypes.POINTER(ctypes.c_double)
    lib.orc_assign.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64, d, d, d,
                               ctypes.c_int]
    lib.orc_measure_pk.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                   ctypes.c_float, ctypes.c_int, d, d, d, d]
    lib.orc_triple_sums.argtypes = [ctypes.c_int64, d, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_int), d, ctypes.c_int]
    ip = ctypes.POINTER(ctypes.c_int)
    lib.orc_count_pairs_exact.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ip, ip, ip,
                                          ctypes.c_int64, ctypes.POINTER(ctypes.c_int64),
                                          ctypes.c_int]
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _lib()
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def nthreads():
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:
        return os.cpu_count() or 1


# --------------------------------------------------------------------------------------------
# derived parameters (pbi4/pbi4.py:86-119)


def itot_count(nbins, kbin, kcenter, iopen, cross=False):
    return len(triangle_list(nbins, kbin, kcenter, iopen, cross))


def triangle_list(nbins, kbin, kcenter, iopen, cross=False):
    """(itot, 3) array of 1-based (l, j, i) in the reference loop order (pbi4.py:105-117)."""
    ishift = int(kcenter / kbin) + (1 if iopen else 0)
    out = []
    for l in range(1, nbins + 1):
        for j in range(1, l + 1):
            imax = min(nbins, j + l - 1 + ishift) if cross else j
            for i in range(max(1, l - j + 1 - ishift), imax + 1):
                out.append((l, j, i))
    return np.array(out, dtype=np.int64).reshape(-1, 3)


def derived_params(params):
    from math import floor
    los = np.array(params["los"], dtype=np.float64)
    d = {"dk": params["kbin"] * 2.0 * np.pi / params["box"],
         "nbins_ps": floor((params["grid"] / 2 + params["kbin"] - params["kcenter"]) / params["kbin"]),
         "nbins_bisp": floor((params["grid"] / 3 + params["kbin"] - params["kcenter"]) / params["kbin"]),
         "kf": 2.0 * np.pi / params["box"],
         "kN": 2 * np.pi / params["box"] * params["grid"] / 2.0,
         "Hz": 100.0 * np.sqrt(params["om"] * (1.0 + params["z"]) ** 3 + params["ol"]),
         "los": los / (los @ los)}
    d["kvec"] = np.array([(i - 1) * d["dk"] + params["kcenter"] * 2.0 * np.pi / params["box"]
                          for i in range(1, d["nbins_ps"] + 1)])
    d["itot"] = itot_count(d["nbins_bisp"], params["kbin"], params["kcenter"], params["iopen"],
                           params.get("cross", False))
    return d


# --------------------------------------------------------------------------------------------
# assignment (powerI4.f90:68-249)


def assign_single_grid(grid, nintp, npart, pos, weight=None, workers=None):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    assert pos.shape == (3, npart)
    w = None if weight is None else np.ascontiguousarray(weight, dtype=np.float64)
    out = np.zeros(grid ** 3, dtype=np.float64)
    lib().orc_assign(grid, nintp, 0, npart, _dp(pos), None if w is None else _dp(w), _dp(out),
                     workers or nthreads())
    return out.reshape((grid, grid, grid), order="F")


def assign_double_grid(grid, nintp, npart, pos, weight=None, workers=None):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    assert pos.shape == (3, npart)
    w = None if weight is None else np.ascontiguousarray(weight, dtype=np.float64)
    out = np.zeros(2 * grid ** 3, dtype=np.float64)
    lib().orc_assign(grid, nintp, 1, npart, _dp(pos), None if w is None else _dp(w), _dp(out),
                     workers or nthreads())
    return out.view(np.complex128).reshape((grid, grid, grid), order="F")


def bspline_weights_closed_form(order, h):
    """Textbook B-spline mass-assignment weights of order 1..4 for a particle at fractional
    offset h from its reference cell (independent of the Fortran expressions)."""
    def M(p, x):  # centred cardinal B-spline of order p evaluated at x
        x = abs(x)
        if p == 1:
            return 1.0 if x < 0.5 else 0.0
        if p == 2:
            return max(0.0, 1.0 - x)
        if p == 3:
            if x < 0.5:
                return 0.75 - x * x
            return 0.5 * (1.5 - x) ** 2 if x < 1.5 else 0.0
        if x < 1.0:
            return (4.0 - 6.0 * x * x + 3.0 * x ** 3) / 6.0
        return (2.0 - x) ** 3 / 6.0 if x < 2.0 else 0.0
    return M(order, h)


# --------------------------------------------------------------------------------------------
# density (pbi4.py:304-341) and fcomb (powerI4.f90:251-378)


def _powi(x, n):
    """x**n for an integer variable exponent, evaluated as libgcc's __powidf2."""
    y = x if n % 2 else np.ones_like(x)
    n >>= 1
    while n:
        x = x * x
        if n % 2:
            y = y * x
        n >>= 1
    return y


def window_table(grid, nintp):
    """tWk(1..N/2+1) of fcomb (powerI4.f90:278-292)."""
    tpiN = np.float64(TWOPI) / np.float64(grid)
    i = np.arange(1, grid // 2 + 2, dtype=np.float64)
    rk = tpiN * (i - 1.0)
    with np.errstate(invalid="ignore", divide="ignore"):
        w = _powi(np.sin(rk / 2.0) / (rk / 2.0), nintp)
    w[0] = 1.0
    return w


def kf3_f32(box):
    """kF3 = twopi*twopi*twopi/box/box/box evaluated in REAL(4) (powerI4.f90:271)."""
    b = np.float32(box)
    return np.float64(np.float32(np.float32(np.float32(np.float32(TWOPI * TWOPI) * TWOPI) / b) / b) / b)


def fcomb(grid, nintp, interlacing, box, fdcl):
    N = grid
    h = N // 2 + 1
    D = np.array(fdcl, dtype=np.complex128)          # dcl = fdcl
    tW = window_table(N, nintp)
    kF3 = kf3_f32(box)
    ix = np.arange(h)
    icx = (N - ix) % N
    if interlacing:
        cf = 1.0 / (4.0 * kF3)
        tpiN = np.float64(TWOPI) / np.float64(N)
        piN = -tpiN / 2.0
        rec = complex(np.cos(piN), np.sin(piN))
        xr = np.empty(h, dtype=np.complex128)
        acc = complex(1.0, 0.0)
        for n in range(h):                            # xrec=xrec*recx, repeated product
            xr[n] = acc
            acc = acc * rec
        X = xr[:, None, None]
        Y = xr[None, :, None]
        Z = xr[None, None, :]
        ci = 1j
        cma = ((ci * X) * Y) * Z
        cmb = ((ci * X) * Y) * np.conj(Z)
        cmc = ((ci * X) * np.conj(Y)) * Z
        cmd = (ci * X) * np.conj(Y * Z)
        Wx, Wy, Wz = tW[:, None, None], tW[None, :, None], tW[None, None, :]
        cfac = cf / ((Wx * Wy) * Wz)
        I, J, K = np.meshgrid(ix, ix, ix, indexing="ij")
        IC, JC, KC = np.meshgrid(icx, icx, icx, indexing="ij")
        c1 = 1.0 + 0.0j
        c000 = D[I, J, K] * (c1 - cma) + np.conj(D[IC, JC, KC]) * (c1 + cma)
        c001 = D[I, J, KC] * (c1 - cmb) + np.conj(D[IC, JC, K]) * (c1 + cmb)
        c010 = D[I, JC, K] * (c1 - cmc) + np.conj(D[IC, J, KC]) * (c1 + cmc)
        c011 = D[I, JC, KC] * (c1 - cmd) + np.conj(D[IC, J, K]) * (c1 + cmd)
        # the eight ordered stores of powerI4.f90:335-342; iterations never share cells
        D[I, J, K] = c000 * cfac
        D[I, J, KC] = c001 * cfac
        D[I, JC, K] = c010 * cfac
        D[I, JC, KC] = c011 * cfac
        D[IC, J, K] = np.conj(D[I, JC, KC])
        D[IC, J, KC] = np.conj(D[I, JC, K])
        D[IC, JC, K] = np.conj(D[I, J, KC])
        D[IC, JC, KC] = np.conj(D[I, J, K])
        return D
    cf = 1.0 / kF3
    iy = np.arange(h)
    icy = (N - iy) % N
    for a in range(h):                                # iz loop
        ica = (N - a) % N
        for b in range(h):                            # iy loop
            icb = (N - b) % N
            cfac = cf / ((tW[:h] * tW[b]) * tW[a])
            xs = np.arange(h)
            D[xs, b, a] = D[xs, b, a] * cfac
            if a != ica:
                D[xs, b, ica] = D[xs, b, ica] * cfac
            if b != icb:
                D[xs, icb, a] = D[xs, icb, a] * cfac
            if a != ica and b != icb:
                D[xs, icb, ica] = D[xs, icb, ica] * cfac
    del icy
    return D


def catalogue_positions(pos, vel, rsd, los, hz, z, box, object_rows=False):
    """IO.read_catalogues' shift, wrap and normalisation, literally on numpy arrays: the
    in-place shift pos[:] += vel * los / Hz * (1 + z) (rsd 1: pbi4/pbi4.py:153-154, 163-164,
    211-212) or pos[:] += vel * los (rsd 2: :179-180), then :241-243.  pos, vel (n, 3);
    object_rows reproduces the Quijote branch, whose positions live in an object array of
    Python floats (:203-207), converted to float64 at the end as f2py does.  Returns (3, n)."""
    pos = np.array(pos, dtype=object) if object_rows else np.array(pos)
    if rsd == 1:
        pos[:] += vel[:] * np.asarray(los) / hz * (1. + z)
    elif rsd == 2:
        pos[:] += vel[:] * np.asarray(los)
    pos = pos.T
    pos = (pos + box) % box
    pos = np.minimum(1. - np.nextafter(0, 1), pos / box)
    return np.asarray(pos, dtype=np.float64) if object_rows else pos


def reduce_array(L, din, grid):
    """powerI4.f90:380-407 literally: the four ordered stores per (iy, iz), vectorised over ix
    (1-based Fortran indices shifted to 0-based).  din (L, L, L) indexed [ix, iy, iz]; returns
    (grid, grid, grid) complex128, zero outside ix <= grid/2 (intent(out) is zero-filled)."""
    din = np.asarray(din)
    dout = np.zeros((grid, grid, grid), dtype=np.complex128)
    h = grid // 2 + 1
    for iz in range(1, grid // 2 + 2):
        icz = (L - iz + 1) % L + 1
        jcz = (grid - iz + 1) % grid + 1
        for iy in range(1, grid // 2 + 2):
            icy = (L - iy + 1) % L + 1
            jcy = (grid - iy + 1) % grid + 1
            dout[:h, iy - 1, iz - 1] = din[:h, iy - 1, iz - 1]
            dout[:h, jcy - 1, iz - 1] = din[:h, icy - 1, iz - 1]
            dout[:h, iy - 1, jcz - 1] = din[:h, iy - 1, icz - 1]
            dout[:h, jcy - 1, jcz - 1] = din[:h, icy - 1, icz - 1]
    return dout


def compute_density(npart, dcl, grid, nintp, interlacing, box, workers=None):
    dcl = dcl / npart                                 # pbi4.py:336
    dcl = scipy.fft.fftn(dcl, workers=workers or nthreads())
    return fcomb(grid, nintp, interlacing, box, dcl)


# --------------------------------------------------------------------------------------------
# multipoles (pbi4.py:343-388, powerI4.f90:409-508)


def _signed_freqs(N):
    i = np.arange(1, N + 1)
    return (np.mod(i + N // 2 - 2, N) - N // 2 + 1).astype(np.float64)


def compute_q2k(grid, box, deltak, i, j):
    N = grid
    kF = np.float64(np.float32(TWOPI / np.float32(box)))
    Q = np.array(deltak, dtype=np.complex128)
    f = kF * _signed_freqs(N)
    h = N // 2 + 1
    v = [f[:h, None, None], f[None, :, None], f[None, None, :]]
    rk2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]
    with np.errstate(invalid="ignore", divide="ignore"):
        R = ((Q[:h] * v[i - 1]) * v[j - 1]) / rk2
    R[0, 0, 0] = 0.0
    Q[:h] = R
    return Q


def compute_q4k(grid, box, deltak, i, j, l, k):
    N = grid
    kF = np.float64(np.float32(TWOPI / np.float32(box)))
    Q = np.array(deltak, dtype=np.complex128)
    f = kF * _signed_freqs(N)
    h = N // 2 + 1
    v = [f[:h, None, None], f[None, :, None], f[None, None, :]]
    rk2 = ((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]) ** 2
    with np.errstate(invalid="ignore", divide="ignore"):
        R = ((((Q[:h] * v[i - 1]) * v[j - 1]) * v[l - 1]) * v[k - 1]) / rk2
    R[0, 0, 0] = 0.0
    Q[:h] = R
    return Q


def compute_density_multipoles(dcl, grid, box, los, workers=None):
    """Literal pbi4.py:359-387 (1 ifftn + 21 fftn)."""
    w = workers or nthreads()
    deltax = scipy.fft.ifftn(dcl, workers=w)
    dcl2 = -0.5 * dcl
    for i in range(3):
        for j in range(i + 1):
            fac = 2.0 if i != j else 1.0
            deltak = scipy.fft.fftn(deltax * los[i] * los[j], workers=w)
            dcl2 = dcl2 + fac * 1.5 * compute_q2k(grid, box, deltak, i + 1, j + 1)
    dcl4 = -7.0 / 8.0 * dcl - 2.5 * dcl2
    for i in range(3):
        for j in range(i + 1):
            for l in range(j + 1):
                for k in range(l + 1):
                    if i == j == l == k:
                        fac = 1
                    if i == j and j == l and l > k:
                        fac = 4
                    if i == j and j > l and l == k:
                        fac = 6
                    if i > j and j == l and l == k:
                        fac = 4
                    if i == j and j > l and l > k:
                        fac = 12
                    if i > j and j == l and l > k:
                        fac = 12
                    if i > j and j > l and l == k:
                        fac = 12
                    deltak = scipy.fft.fftn(deltax * los[i] * los[j] * los[l] * los[k], workers=w)
                    dcl4 = dcl4 + fac * (35.0 / 8.0) * compute_q4k(grid, box, deltak,
                                                                  i + 1, j + 1, l + 1, k + 1)
    return dcl2, dcl4


def legendre_factors(grid, los):
    """mu^2 = (l.k)^2/k^2 on the ix <= N/2 half (los already divided by |los|^2 as in
    pbi4.py:103), and L2, L4 with the DC convention of compute_q2k/q4k (Q = 0 at k = 0)."""
    N = grid
    f = _signed_freqs(N)
    h = N // 2 + 1
    kx, ky, kz = f[:h, None, None], f[None, :, None], f[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    with np.errstate(invalid="ignore", divide="ignore"):
        mu2 = (los[0] * kx + los[1] * ky + los[2] * kz) ** 2 / k2
    mu2[0, 0, 0] = 0.0
    L2 = -0.5 + 1.5 * mu2
    L4 = 0.375 - 3.75 * mu2 + 4.375 * mu2 * mu2
    return L2, L4


# --------------------------------------------------------------------------------------------
# power spectrum binning (powerI4.f90:510-693)


def _fortran_complex(a):
    a = np.asarray(a, dtype=np.complex128)
    return np.asfortranarray(a).ravel(order="F").view(np.float64)


def measure_pk(grid, box, kbin, kcenter, nbins, dcl1, dcl2):
    a, b = _fortran_complex(dcl1), _fortran_complex(dcl2)
    pout = np.zeros(5 * nbins)
    lib().orc_measure_pk(0, grid, box, kbin, kcenter, nbins, _dp(a), _dp(b), _dp(b), _dp(pout))
    return pout.reshape((5, nbins), order="F")


def measure_pk_multipoles(grid, box, kbin, kcenter, nbins, dcl, dcl2, dcl4):
    a, b, c = _fortran_complex(dcl), _fortran_complex(dcl2), _fortran_complex(dcl4)
    pout = np.zeros(5 * nbins)
    lib().orc_measure_pk(1, grid, box, kbin, kcenter, nbins, _dp(a), _dp(b), _dp(c), _dp(pout))
    return pout.reshape((5, nbins), order="F")


def pk_bin_index(grid, box, kbin, kcenter, x0=0, x1=None):
    """imk of measure_pk for every half-space mode (ix, iy, iz) 0-based, 0 = not selected
    (powerI4.f90:548-569).  Used to check mode counts against ps.dat.  x0, x1: only the planes
    x0 <= ix < x1 (default all N/2 + 1), for bounded memory at large N."""
    N = grid
    kF = np.float64(np.float32(TWOPI / np.float32(box)))
    f = _signed_freqs(N)
    h = N // 2 + 1
    x1 = h if x1 is None else min(int(x1), h)
    fxs = f[x0:x1]
    rkx = (kF * fxs)[:, None, None]
    rky = (kF * f)[None, :, None]
    rkz = (kF * f)[None, None, :]
    rk = np.sqrt(rkx ** 2 + rky ** 2 + rkz ** 2)
    dk = np.float64(np.float32(kbin)) * kF
    kcb = np.float64(np.float32(kcenter) / np.float32(kbin))
    imk = fnint(rk / dk - kcb - TINY) + 1
    fx, fy, fz = fxs[:, None, None], f[None, :, None], f[None, None, :]
    sel = (fx != 0) | ((fx == 0) & (fy > 0)) | ((fx == 0) & (fy == 0) & (fz > 0))
    return np.where(sel, imk, 0)


def pk_mode_counts(grid, box, kbin, kcenter, nbins, planes=32):
    """Mode count per P(k) bin (the co column of measure_pk), as a bincount of pk_bin_index
    taken over blocks of x-planes."""
    h = grid // 2 + 1
    co = np.zeros(nbins + 2, dtype=np.int64)
    for x0 in range(0, h, planes):
        imk = pk_bin_index(grid, box, kbin, kcenter, x0, x0 + planes)
        imk = np.where((imk >= 1) & (imk <= nbins), imk, 0)
        co += np.bincount(imk.ravel(), minlength=nbins + 2)[:nbins + 2]
    return co[1:nbins + 1]


# --------------------------------------------------------------------------------------------
# bispectrum (bispectrumI4.f90)


def shell_of_d2(d2, kbin, kcenter):
    """idist = nint(sqrt(d2)/kbin - kcenter/kbin - tiny) + 1 (bispectrumI4.f90:73)."""
    kb = np.float64(np.float32(kbin))
    kcb = np.float64(np.float32(kcenter) / np.float32(kbin))
    return fnint(np.sqrt(np.asarray(d2, dtype=np.float64)) / kb - kcb - TINY) + 1


def shell_index_grid(grid, kbin, kcenter):
    """idist for every full-grid point, array indexed [i, j, l] like the reference loops."""
    N = grid
    c = np.arange(N)
    d = np.minimum(c, N - c).astype(np.float64)
    d2 = (d[:, None, None] * d[:, None, None] + d[None, :, None] * d[None, :, None]) \
        + d[None, None, :] * d[None, None, :]
    return shell_of_d2(d2, kbin, kcenter)


def full_spectrum(dcl):
    """Hermitian rebuild of bispectrumI4.f90:116-136 (indexed [i, j, l] = dcl(i+1, j+1, l+1))."""
    D = np.asarray(dcl, dtype=np.complex128)
    N = D.shape[0]
    h = N // 2 + 1
    F = np.empty_like(D)
    F[:h] = D[:h]
    i = np.arange(h, N)
    ic = (N - i) % N
    jc = (N - np.arange(N)) % N
    F[h:] = np.conj(D[ic][:, jc][:, :, jc])
    sp = [0, N // 2]
    for a in sp:
        for b in sp:
            for c in sp:
                F[a, b, c] = F[a, b, c].real
    return F


def shell_maps(field, grid, nbins, kbin, kcenter, workers=None, shells=None):
    """Per-shell real maps M_s(x) = Re FFTW_FORWARD(masked spectrum) (bispectrumI4.f90:142-157),
    returned as (nbins, N^3) FP64, with nmodes per shell.  shells: optional 1-based list; then
    only those maps are built, in that order, with their nmodes."""
    idist = shell_index_grid(grid, kbin, kcenter)
    F = full_spectrum(field)
    which = list(range(1, nbins + 1)) if shells is None else [int(s) for s in shells]
    maps = np.empty((len(which), grid ** 3), dtype=np.float64)
    nmodes = np.zeros(len(which), dtype=np.int64)
    for k, s in enumerate(which):
        m = idist == s
        nmodes[k] = int(m.sum())
        maps[k] = scipy.fft.fftn(np.where(m, F, 0.0), workers=workers or nthreads()).real.ravel()
    return maps, nmodes


def kf3_bk(kf):
    """kf3 = kf**3 with kf REAL(4) (bispectrumI4.f90:46)."""
    k = np.float32(kf)
    return np.float64(np.float32(np.float32(k * k) * k))


def _triangle_map_index(nbins, kbin, kcenter, iopen, cross=False):
    tl = triangle_list(nbins, kbin, kcenter, iopen, cross)
    return tl, np.ascontiguousarray(np.stack([tl[:, 2] - 1, tl[:, 1] - 1, tl[:, 0] - 1], axis=1),
                                    dtype=np.int32)


def triple_sums(maps_a, maps_c, idx, workers=None):
    """sum_n A[i] A[j] C[l] for each idx row (i, j, l); A and C stacked so one C call does it."""
    nb = maps_a.shape[0]
    stack = np.ascontiguousarray(np.concatenate([maps_a, maps_c], axis=0))
    tri = idx.copy()
    tri[:, 2] += nb
    out = np.zeros(len(tri))
    lib().orc_triple_sums(maps_a.shape[1], _dp(stack), len(tri), _ip(np.ascontiguousarray(tri)),
                          _dp(out), workers or nthreads())
    return out


def measure_bk(nbins, grid, kbin, kcenter, iopen, kf, itot, npart, counts, dcl,
               dcl2=None, dcl4=None, workers=None, maps=None):
    """measure_bk (9 rows) or, when dcl2/dcl4 are given, measure_bk_multipoles (11 rows)."""
    counts = np.asarray(counts, dtype=np.float32).astype(np.float64)   # f2py REAL(4) coercion
    tl, idx = _triangle_map_index(nbins, kbin, kcenter, iopen)
    assert len(tl) == itot
    kb = np.float64(np.float32(kbin))
    kc = np.float64(np.float32(kcenter))
    tk = kc + np.arange(nbins, dtype=np.float64) * kb
    kf3 = kf3_bk(kf)
    powSN = 1.0 / np.float64(npart) / kf3
    powSN2 = powSN * powSN
    gridD = np.float64(grid) ** 3
    if maps is None:
        maps, nmodes = shell_maps(dcl, grid, nbins, kbin, kcenter, workers)
    else:
        maps, nmodes = maps
    pw = kf3 * np.sum(maps * maps, axis=1) / gridD / nmodes.astype(np.float64) - powSN
    s0 = triple_sums(maps, maps, idx, workers)
    rows = 9 if dcl2 is None else 11
    bout = np.zeros((rows, itot))
    i, j, l = idx[:, 0], idx[:, 1], idx[:, 2]
    bispSN = ((pw[i] + pw[j]) + pw[l]) * powSN + powSN2
    bout[0], bout[1], bout[2] = tk[l], tk[j], tk[i]
    bout[3], bout[4], bout[5] = pw[l], pw[j], pw[i]
    bout[6] = kf3 * s0 / counts + bispSN
    bout[7] = bispSN
    bout[8] = (counts / gridD).astype(np.float32).astype(np.float64)
    if dcl2 is not None:
        m2, _ = shell_maps(dcl2, grid, nbins, kbin, kcenter, workers)
        s2 = triple_sums(maps, m2, idx, workers)
        del m2
        m4, _ = shell_maps(dcl4, grid, nbins, kbin, kcenter, workers)
        s4 = triple_sums(maps, m4, idx, workers)
        bout[9] = 5 * kf3 * s2 / counts
        bout[10] = 9 * kf3 * s4 / counts
    return bout


def measure_bk_subset(nbins, grid, kbin, kcenter, iopen, kf, itot, npart, counts, dcl, rows_m,
                      dcl2=None, dcl4=None, workers=None):
    """Columns rows_m (0-based triangle indices of the reference loop) of measure_bk /
    measure_bk_multipoles, with shell maps built only for the shells those triangles use.  Each
    column is computed exactly as measure_bk computes it (same maps, same per-triangle sums, same
    row formulas), so it equals measure_bk(...)[:, rows_m] bit for bit; counts has all itot."""
    counts = np.asarray(counts, dtype=np.float32).astype(np.float64)
    tl, idx = _triangle_map_index(nbins, kbin, kcenter, iopen)
    assert len(tl) == itot
    rows_m = np.asarray(rows_m, dtype=np.int64)
    used = sorted(set(idx[rows_m].ravel().tolist()))           # 0-based shells
    pos_of = {s: k for k, s in enumerate(used)}
    sub = np.ascontiguousarray(np.vectorize(pos_of.get)(idx[rows_m]).reshape(-1, 3), dtype=np.int32)
    kb = np.float64(np.float32(kbin))
    kc = np.float64(np.float32(kcenter))
    tk = kc + np.arange(nbins, dtype=np.float64) * kb
    kf3 = kf3_bk(kf)
    powSN = 1.0 / np.float64(npart) / kf3
    powSN2 = powSN * powSN
    gridD = np.float64(grid) ** 3
    maps, nmodes = shell_maps(dcl, grid, nbins, kbin, kcenter, workers, [s + 1 for s in used])
    pw = kf3 * np.sum(maps * maps, axis=1) / gridD / nmodes.astype(np.float64) - powSN
    s0 = triple_sums(maps, maps, sub, workers)
    cm = counts[rows_m]
    rows = 9 if dcl2 is None else 11
    bout = np.zeros((rows, len(rows_m)))
    i, j, l = idx[rows_m, 0], idx[rows_m, 1], idx[rows_m, 2]
    si, sj, sl = sub[:, 0], sub[:, 1], sub[:, 2]
    bispSN = ((pw[si] + pw[sj]) + pw[sl]) * powSN + powSN2
    bout[0], bout[1], bout[2] = tk[l], tk[j], tk[i]
    bout[3], bout[4], bout[5] = pw[sl], pw[sj], pw[si]
    bout[6] = kf3 * s0 / cm + bispSN
    bout[7] = bispSN
    bout[8] = (cm / gridD).astype(np.float32).astype(np.float64)
    if dcl2 is not None:
        m2, _ = shell_maps(dcl2, grid, nbins, kbin, kcenter, workers, [s + 1 for s in used])
        s2 = triple_sums(maps, m2, sub, workers)
        del m2
        m4, _ = shell_maps(dcl4, grid, nbins, kbin, kcenter, workers, [s + 1 for s in used])
        s4 = triple_sums(maps, m4, sub, workers)
        bout[9] = 5 * kf3 * s2 / cm
        bout[10] = 9 * kf3 * s4 / cm
    return bout


def measure_bk_cross(nbins, grid, kbin, kcenter, iopen, kf, itot, npart, counts, dcl1, dcl2,
                     workers=None):
    """measure_bk_cross (bispectrumI4.f90:518-761): rows k_l, k_j, k_i, kf3 sum_x B_i A_j A_l / N_tr
    and N_tr / N^3 (:751-755), with A the shell maps of dcl1 (:644-657), B those of dcl2
    (:691-703), over the cross loop's wider range of i (:733-735).  No shot noise and no power
    columns; npart is unused, as in the reference."""
    counts = np.asarray(counts, dtype=np.float32).astype(np.float64)   # f2py REAL(4) coercion
    tl = triangle_list(nbins, kbin, kcenter, iopen, True)
    assert len(tl) == itot
    kb = np.float64(np.float32(kbin))
    kc = np.float64(np.float32(kcenter))
    tk = kc + np.arange(nbins, dtype=np.float64) * kb
    kf3 = kf3_bk(kf)
    gridD = np.float64(grid) ** 3
    A, _ = shell_maps(dcl1, grid, nbins, kbin, kcenter, workers)
    B, _ = shell_maps(dcl2, grid, nbins, kbin, kcenter, workers)
    l, j, i = tl[:, 0] - 1, tl[:, 1] - 1, tl[:, 2] - 1
    idx = np.ascontiguousarray(np.stack([j, l, i], axis=1), dtype=np.int32)
    s = triple_sums(A, B, idx, workers)
    bout = np.zeros((5, itot))
    bout[0], bout[1], bout[2] = tk[l], tk[j], tk[i]
    bout[3] = kf3 * s / counts
    bout[4] = (counts / gridD).astype(np.float32).astype(np.float64)
    return bout


def count_triangles(nbins, grid, kbin, kcenter, iopen, cross=False, workers=None):
    """FFT form of count_triangles (bispectrumI4.f90:808-908), FP64."""
    idist = shell_index_grid(grid, kbin, kcenter)
    maps = np.empty((nbins, grid ** 3))
    for s in range(1, nbins + 1):
        maps[s - 1] = scipy.fft.fftn((idist == s).astype(np.complex128),
                                     workers=workers or nthreads()).real.ravel()
    tl, idx = _triangle_map_index(nbins, kbin, kcenter, iopen, cross)
    return triple_sums(maps, maps, idx, workers)


def shell_modes(grid, nbins, kbin, kcenter):
    """Signed components (-N/2, N/2] of every grid mode in shells 1..nbins, sorted by shell."""
    N = grid
    idist = shell_index_grid(grid, kbin, kcenter)
    c = np.arange(N)
    s = np.where(c > N // 2, c - N, c)
    I, J, L = np.meshgrid(s, s, s, indexing="ij")
    sel = (idist >= 1) & (idist <= nbins)
    sh = idist[sel]
    order = np.argsort(sh, kind="stable")
    modes = np.stack([I[sel], J[sel], L[sel]], axis=1)[order].astype(np.int32)
    return modes, sh[order].astype(np.int32)


def count_triangles_exact(nbins, grid, kbin, kcenter, iopen, cross=False, workers=None):
    """Exact integer N_tr for the triangle list: #{q1 in S_i, q2 in S_j, q3 in S_l :
    q1+q2+q3 = 0 mod N} (the integer meaning of count_triangles), by brute force over pairs."""
    modes, msh = shell_modes(grid, nbins, kbin, kcenter)
    N = grid
    maxd = N // 2
    max_d2 = 3 * maxd * maxd
    table = shell_of_d2(np.arange(max_d2 + 1), kbin, kcenter)
    table = np.where((table >= 1) & (table <= nbins), table, 0).astype(np.int32)
    S = nbins + 1
    T = np.zeros(S ** 3, dtype=np.int64)
    lib().orc_count_pairs_exact(grid, nbins, len(modes), _ip(np.ascontiguousarray(modes)),
                                _ip(np.ascontiguousarray(msh)), _ip(table), max_d2,
                                T.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                workers or nthreads())
    T = T.reshape(S, S, S)
    tl = triangle_list(nbins, kbin, kcenter, iopen, cross)
    out = np.empty(len(tl), dtype=np.int64)
    for m, (l, j, i) in enumerate(tl):
        a, b = min(i, j), max(i, j)
        out[m] = T[a, b, l]
    return out
