"""Device-level API over the C ABI: torch tensors in, torch tensors out, asynchronous on the
current (or given) HIP stream.  Each method maps to one entry point of include/pbi4_amd.h.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import AssignDesc, BkDesc, DensityDesc, Pbi4Error, PkDesc, lib

ORDER_NAMES = {1: "NGP", 2: "CIC", 3: "TSC", 4: "PCS"}


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(stream):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _check_dev(t, name, dtype=None):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise Pbi4Error(f"{name} must be a device tensor")
    if dtype is not None and t.dtype != dtype:
        raise Pbi4Error(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise Pbi4Error(f"{name} must be contiguous")


def kf3_f32(kf):
    """kf**3 in REAL(4) arithmetic (bispectrumI4.f90:46)."""
    k = np.float32(kf)
    return float(np.float32(np.float32(k * k) * k))


def _assign_ws_bytes(desc, npart):
    n = lib().pbi4_assign_workspace(ctypes.byref(desc), int(npart))
    if n == 0:
        raise Pbi4Error("pbi4_assign_workspace: " + lib().pbi4_last_error(None).decode())
    return n


def assign_workspace_bytes(grid, order, interlacing, npart):
    """Workspace bytes pbi4_assign needs for these sizes (host only, no GPU).  Raises Pbi4Error
    with the library's reason for a refused shape, e.g. a grid with no tile edge in [4, 16]."""
    d = AssignDesc(int(grid), int(order), int(bool(interlacing)), _lib.POS_BOX, 1.0,
                   (ctypes.c_float * 3)(0.0, 0.0, 0.0))
    return _assign_ws_bytes(d, npart)


class Context:
    """One pbi4_ctx (rocFFT plans and integer tables are cached in it)."""

    def __init__(self, device=0):
        if not torch.cuda.is_available():
            raise Pbi4Error("no GPU visible: the pbi4_amd path runs only on the GPU")
        self.device = torch.device("cuda", device)
        torch.cuda.init()
        h = ctypes.c_void_p()
        self._check(lib().pbi4_ctx_create(device, ctypes.byref(h)), None)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().pbi4_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, h="self"):
        if rc != _lib.PBI4_OK:
            msg = lib().pbi4_last_error(self.h if h == "self" else None)
            raise Pbi4Error(f"pbi4 error {rc}: {msg.decode() if msg else ''}")

    def _ws(self, nbytes):
        return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=self.device)

    # ------------------------------------------------------------- assignment
    def assign_workspace(self, grid, order, interlacing, npart):
        d = AssignDesc(int(grid), int(order), int(bool(interlacing)), _lib.POS_BOX, 1.0,
                       (ctypes.c_float * 3)(0.0, 0.0, 0.0))
        return _assign_ws_bytes(d, npart)

    def density_workspace(self, grid, order, interlacing, box, npart):
        d = DensityDesc(int(grid), int(order), int(bool(interlacing)), float(box), float(npart))
        return lib().pbi4_density_workspace(self.h, ctypes.byref(d))

    def _ws_or(self, ws, nbytes, name):
        if ws is None:
            return self._ws(nbytes)
        _check_dev(ws, name, torch.uint8)
        if ws.numel() < nbytes:
            raise Pbi4Error(f"{name}: {ws.numel()} bytes, {nbytes} needed")
        return ws

    def assign(self, pos, grid, order, interlacing=True, vel=None, box=1.0, rsd=None,
               normalised=True, out=None, ws=None, stream=None):
        """Grids (1 or 2, N, N, N) float32 of the summed weights (powerI4.f90:68-249).
        ws: optional caller-owned uint8 device workspace (else one is allocated)."""
        _check_dev(pos, "pos", torch.float32)
        if pos.dim() != 2 or pos.shape[0] != 3:
            raise Pbi4Error("pos must have shape (3, npart)")
        npart = pos.shape[1]
        if vel is not None:
            _check_dev(vel, "vel", torch.float32)
            if vel.shape != pos.shape:
                raise Pbi4Error("vel must have the shape of pos")
        d = AssignDesc(int(grid), int(order), int(bool(interlacing)),
                       _lib.POS_NORMALISED if normalised else _lib.POS_BOX, float(box),
                       (ctypes.c_float * 3)(*(rsd if rsd is not None else (0.0, 0.0, 0.0))))
        ng = 2 if interlacing else 1
        if out is None:
            out = torch.empty((ng, grid, grid, grid), dtype=torch.float32, device=self.device)
        ws = self._ws_or(ws, _assign_ws_bytes(d, npart), "ws")
        self._check(lib().pbi4_assign(self.h, ctypes.byref(d), _ptr(pos), _ptr(vel), npart,
                                      _ptr(out), _ptr(ws), ws.numel(), _stream(stream)))
        return out

    # ------------------------------------------------------------- density
    def density(self, grids, grid, order, interlacing, box, npart, out=None, ws=None, stream=None):
        """Half spectrum (N, N, N/2+1) complex64 of compute_density + fcomb."""
        _check_dev(grids, "grids", torch.float32)
        d = DensityDesc(int(grid), int(order), int(bool(interlacing)), float(box), float(npart))
        if out is None:
            out = torch.empty((grid, grid, grid // 2 + 1), dtype=torch.complex64, device=self.device)
        ws = self._ws_or(ws, lib().pbi4_density_workspace(self.h, ctypes.byref(d)), "ws")
        self._check(lib().pbi4_compute_density(self.h, ctypes.byref(d), _ptr(grids), _ptr(out),
                                               _ptr(ws), ws.numel(), _stream(stream)))
        return out

    def assign_density(self, pos, grid, order, interlacing, box, npart, vel=None, rsd=None,
                       normalised=True, grids=None, out=None, ws=None, stream=None):
        """assign followed by density in one call (pbi4_assign_density): the half spectrum
        (N, N, N/2+1) complex64.  grids: optional (ng, N, N, N) float32 scratch (its contents
        are undefined afterwards); ws: optional uint8 workspace of at least
        max(assign_workspace, density_workspace) bytes."""
        _check_dev(pos, "pos", torch.float32)
        if pos.dim() != 2 or pos.shape[0] != 3:
            raise Pbi4Error("pos must have shape (3, npart)")
        n = pos.shape[1]
        if vel is not None:
            _check_dev(vel, "vel", torch.float32)
            if vel.shape != pos.shape:
                raise Pbi4Error("vel must have the shape of pos")
        ad = AssignDesc(int(grid), int(order), int(bool(interlacing)),
                        _lib.POS_NORMALISED if normalised else _lib.POS_BOX, float(box),
                        (ctypes.c_float * 3)(*(rsd if rsd is not None else (0.0, 0.0, 0.0))))
        dd = DensityDesc(int(grid), int(order), int(bool(interlacing)), float(box), float(npart))
        ng = 2 if interlacing else 1
        if grids is None:
            grids = torch.empty((ng, grid, grid, grid), dtype=torch.float32, device=self.device)
        _check_dev(grids, "grids", torch.float32)
        if grids.numel() < ng * grid ** 3:
            raise Pbi4Error("grids holds fewer than ng * N^3 floats")
        if out is None:
            out = torch.empty((grid, grid, grid // 2 + 1), dtype=torch.complex64, device=self.device)
        need = max(_assign_ws_bytes(ad, n), lib().pbi4_density_workspace(self.h, ctypes.byref(dd)))
        ws = self._ws_or(ws, need, "ws")
        self._check(lib().pbi4_assign_density(self.h, ctypes.byref(ad), ctypes.byref(dd), _ptr(pos),
                                              _ptr(vel), n, _ptr(grids), _ptr(out), _ptr(ws),
                                              ws.numel(), _stream(stream)))
        return out

    # ------------------------------------------------------------- full-array forms
    def fcomb_full(self, fdcl, grid, nintp, interlacing, box, out=None, stream=None):
        """powerI4.fcomb on a full (N, N, N) complex64 [z][y][x] array."""
        _check_dev(fdcl, "fdcl", torch.complex64)
        if out is None:
            out = torch.empty_like(fdcl)
        self._check(lib().pbi4_fcomb(self.h, int(grid), int(nintp), int(bool(interlacing)),
                                     float(box), _ptr(fdcl), _ptr(out), _stream(stream)))
        return out

    def compute_qk(self, deltak, grid, idx, out=None, stream=None):
        """compute_q2k (len(idx) == 2) / compute_q4k (len(idx) == 4), 1-based axes."""
        _check_dev(deltak, "deltak", torch.complex64)
        if out is None:
            out = torch.empty_like(deltak)
        arr = (ctypes.c_int * 4)(*(list(idx) + [1] * (4 - len(idx))))
        self._check(lib().pbi4_compute_qk(self.h, int(grid), len(idx), arr, _ptr(deltak), _ptr(out),
                                          _stream(stream)))
        return out

    def catalogue_positions(self, pos, vel=None, rsd=_lib.RSD_NONE, los=(0.0, 0.0, 1.0), hz=1.0,
                            zf=1.0, box=1000.0, out_dtype=None, out=None, stream=None):
        """Redshift-space shift, wrap and normalisation of IO.read_catalogues on the device:
        pos, vel (n, 3) float32/float64 rows -> (3, n) positions in [0, 1] of out_dtype (default
        pos's dtype); see pbi4_catalogue_positions for the arithmetic."""
        _check_dev(pos, "pos")
        if pos.dim() != 2 or pos.shape[1] != 3 or pos.dtype not in (torch.float32, torch.float64):
            raise Pbi4Error("pos must be an (n, 3) float32 or float64 tensor")
        if vel is not None:
            _check_dev(vel, "vel")
            if vel.shape != pos.shape or vel.dtype not in (torch.float32, torch.float64):
                raise Pbi4Error("vel must be float32 or float64 with pos's shape")
        if rsd != _lib.RSD_NONE and vel is None:
            raise Pbi4Error("a redshift-space shift needs velocities")
        out_dtype = pos.dtype if out_dtype is None else out_dtype
        n = pos.shape[0]
        if out is None:
            out = torch.empty((3, n), dtype=out_dtype, device=pos.device)
        _check_dev(out, "out", out_dtype)
        l3 = (ctypes.c_double * 3)(*[float(x) for x in los])
        f64 = torch.float64
        self._check(lib().pbi4_catalogue_positions(
            self.h, _ptr(pos), int(pos.dtype == f64), _ptr(vel), int(vel is not None and vel.dtype == f64),
            int(n), int(rsd), l3, float(hz), float(zf), float(box), _ptr(out), int(out_dtype == f64),
            _stream(stream)))
        return out

    def reduce_array(self, din, L, N, nx=None, out_dtype=None, out=None, stream=None):
        """powerI4.reduce_array on the device: din a [z][y][x] (L, L, lx) complex tensor with
        lx >= N/2+1, result (N, N, nx) of out_dtype (default din's dtype; complex128 ->
        complex64 rounds) with x > N/2 zero."""
        _check_dev(din, "din")
        if din.dtype not in (torch.complex64, torch.complex128) or din.dim() != 3:
            raise Pbi4Error("din must be a 3-D complex64 or complex128 tensor")
        if din.shape[0] != L or din.shape[1] != L:
            raise Pbi4Error(f"din must be ({L}, {L}, lx), got {tuple(din.shape)}")
        nx = N // 2 + 1 if nx is None else int(nx)
        out_dtype = din.dtype if out_dtype is None else out_dtype
        if out is None:
            out = torch.empty((N, N, nx), dtype=out_dtype, device=din.device)
        _check_dev(out, "out", out_dtype)
        nbytes = {torch.complex64: 8, torch.complex128: 16}
        self._check(lib().pbi4_reduce_array(self.h, int(L), int(din.shape[2]), int(N), nx,
                                            nbytes[din.dtype], nbytes[out_dtype], _ptr(din),
                                            _ptr(out), _stream(stream)))
        return out

    # ------------------------------------------------------------- P(k)
    def pk(self, delta, grid, box, kbin, kcenter, nbins, multipoles=False, los=(0.0, 0.0, 1.0),
           delta_b=None, delta2=None, delta4=None, out=None, stream=None):
        """pout as a (nbins, 5) float64 tensor (the reference's Fortran (5, nbins)).
        delta_b: second field of a real-space cross spectrum; delta2/delta4: explicit fields."""
        _check_dev(delta, "delta", torch.complex64)
        for t, nm in ((delta_b, "delta_b"), (delta2, "delta2"), (delta4, "delta4")):
            if t is not None:
                _check_dev(t, nm, torch.complex64)
        d = PkDesc(int(grid), float(box), float(kbin), float(kcenter), int(nbins),
                   int(bool(multipoles)), (ctypes.c_double * 3)(*[float(x) for x in los]))
        if out is None:
            out = torch.empty((nbins, 5), dtype=torch.float64, device=self.device)
        ws = self._ws(lib().pbi4_pk_workspace(self.h, ctypes.byref(d)))
        self._check(lib().pbi4_measure_pk(self.h, ctypes.byref(d), _ptr(delta), _ptr(delta_b),
                                          _ptr(delta2), _ptr(delta4), _ptr(out), _ptr(ws),
                                          ws.numel(), _stream(stream)))
        return out

    # ------------------------------------------------------------- triangle counts
    def count_triangles(self, nbins, grid, kbin, kcenter, iopen, cross, itot):
        """Exact N_tr (int64 numpy array) in the reference loop order."""
        out = np.zeros(int(itot), dtype=np.int64)
        self._check(lib().pbi4_count_triangles(
            self.h, int(nbins), int(grid), float(kbin), float(kcenter), int(bool(iopen)),
            int(bool(cross)), int(itot), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return out

    # ------------------------------------------------------------- bispectrum
    def bk_desc(self, nbins, grid, kbin, kcenter, iopen, kf, itot, npart, multipoles=False,
                los=(0.0, 0.0, 1.0), precision=_lib.PREC_BF16X3, maps_grid=0):
        return BkDesc(int(nbins), int(grid), float(kbin), float(kcenter), int(bool(iopen)),
                      float(kf), int(itot), int(npart), int(bool(multipoles)),
                      (ctypes.c_double * 3)(*[float(x) for x in los]), int(precision),
                      int(maps_grid))

    def maps_grid(self, desc):
        r = lib().pbi4_bk_maps_grid(self.h, ctypes.byref(desc))
        if r < 0:
            self._check(r)
        return r

    def bk_workspace(self, desc):
        n = lib().pbi4_bk_workspace(self.h, ctypes.byref(desc))
        if n == 0:
            raise Pbi4Error("pbi4_bk_workspace: " + lib().pbi4_last_error(self.h).decode())
        return n

    def bk(self, delta, counts, desc, delta2=None, delta4=None, out=None, ws=None, stream=None):
        """bout as an (itot, 9 or 11) float64 tensor (the reference's Fortran (rows, itot)).
        counts: the reference's REAL(4) counts(itot) as a float32 device tensor."""
        _check_dev(delta, "delta", torch.complex64)
        _check_dev(counts, "counts", torch.float32)
        if counts.numel() != desc.itot:
            raise Pbi4Error("counts must have itot entries")
        rows = 11 if desc.multipoles else 9
        if out is None:
            out = torch.empty((desc.itot, rows), dtype=torch.float64, device=self.device)
        if ws is None:
            ws = self._ws(self.bk_workspace(desc))
        self._check(lib().pbi4_measure_bk(self.h, ctypes.byref(desc), _ptr(counts), _ptr(delta),
                                          _ptr(delta2), _ptr(delta4), _ptr(out), _ptr(ws),
                                          ws.numel(), _stream(stream)))
        return out

    def bk_cross_workspace(self, desc):
        n = lib().pbi4_bk_cross_workspace(self.h, ctypes.byref(desc))
        if n == 0:
            raise Pbi4Error("pbi4_bk_cross_workspace: " + lib().pbi4_last_error(self.h).decode())
        return n

    def bk_cross(self, delta1, delta2, counts, desc, out=None, ws=None, stream=None):
        """Cross bispectrum (bispectrumI4.f90:518-761) as an (itot, 5) float64 tensor: k_l, k_j,
        k_i, B, N_tr/N^3.  counts: the cross loop's REAL(4) counts as a float32 device tensor."""
        _check_dev(delta1, "delta1", torch.complex64)
        _check_dev(delta2, "delta2", torch.complex64)
        _check_dev(counts, "counts", torch.float32)
        if counts.numel() != desc.itot:
            raise Pbi4Error("counts must have itot entries")
        if out is None:
            out = torch.empty((desc.itot, 5), dtype=torch.float64, device=self.device)
        if ws is None:
            ws = self._ws(self.bk_cross_workspace(desc))
        self._check(lib().pbi4_measure_bk_cross(self.h, ctypes.byref(desc), _ptr(counts),
                                                _ptr(delta1), _ptr(delta2), _ptr(out), _ptr(ws),
                                                ws.numel(), _stream(stream)))
        return out

    # ------------------------------------------------------------- fused P_l + B_l
    def pk_bk_workspace(self, pk_desc, bk_desc):
        n = lib().pbi4_pk_bk_workspace(self.h, ctypes.byref(pk_desc), ctypes.byref(bk_desc))
        if n == 0:
            raise Pbi4Error("pbi4_pk_bk_workspace: " + lib().pbi4_last_error(self.h).decode())
        return n

    def pk_desc(self, grid, box, kbin, kcenter, nbins, multipoles=False, los=(0.0, 0.0, 1.0)):
        return PkDesc(int(grid), float(box), float(kbin), float(kcenter), int(nbins),
                      int(bool(multipoles)), (ctypes.c_double * 3)(*[float(x) for x in los]))

    def pk_bk(self, delta, counts, pk_desc, bk_desc, pout=None, bout=None, ws=None, stream=None):
        """measure_pk + measure_bk on one delta with a single binning pass (serving path)."""
        _check_dev(delta, "delta", torch.complex64)
        _check_dev(counts, "counts", torch.float32)
        rows = 11 if bk_desc.multipoles else 9
        if pout is None:
            pout = torch.empty((pk_desc.nbins, 5), dtype=torch.float64, device=self.device)
        if bout is None:
            bout = torch.empty((bk_desc.itot, rows), dtype=torch.float64, device=self.device)
        ws = self._ws_or(ws, self.pk_bk_workspace(pk_desc, bk_desc), "ws")
        self._check(lib().pbi4_measure_pk_bk(self.h, ctypes.byref(pk_desc), ctypes.byref(bk_desc),
                                             _ptr(counts), _ptr(delta), _ptr(pout), _ptr(bout),
                                             _ptr(ws), ws.numel(), _stream(stream)))
        return pout, bout

    # ------------------------------------------------------------- triangle sums
    def triangle_sums(self, maps, nb, nell, K, precision, stream=None):
        """C[(i<=j)][(ell,l)] (float64 tensor) for maps (nb*nell, Kstride) float32."""
        _check_dev(maps, "maps", torch.float32)
        Kstride = maps.shape[1]
        C = torch.empty((nb * (nb + 1) // 2, nb * nell), dtype=torch.float64, device=self.device)
        n = lib().pbi4_triangle_sums_workspace(nb, nell, K, Kstride, precision)
        if n == 0:
            raise Pbi4Error("pbi4_triangle_sums_workspace: " + lib().pbi4_last_error(None).decode())
        ws = self._ws(n)
        self._check(lib().pbi4_triangle_sums(self.h, nb, nell, K, Kstride, precision, _ptr(maps),
                                             _ptr(C), _ptr(ws), ws.numel(), _stream(stream)))
        return C

    # ------------------------------------------------------------- timing
    def timing_enable(self, kernel_name):
        self._check(lib().pbi4_timing_enable(self.h, kernel_name.encode() if kernel_name else None))

    def timing_read(self):
        t = ctypes.c_double()
        n = ctypes.c_int64()
        self._check(lib().pbi4_timing_read(self.h, ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value


_DEFAULT = {}


def default_context(device=0):
    if device not in _DEFAULT:
        _DEFAULT[device] = Context(device)
    return _DEFAULT[device]
