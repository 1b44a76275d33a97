"""Drop-in for the reference's pbi4/pbi4.py: IO, DENSITIES and MEASUREMENTS with the same method
names, arguments, config keys and float32 record-array outputs, running on the GPU.

    import pbi4_amd as pb            # instead of `import pbi4 as pb`
    io = pb.IO(config_file='config.json', verbose=True, scriptmode=True)
    ...                              # the call sequence of Tutorials/Script/test.py unchanged

Differences a user can see:
  * DENSITIES returns device-resident objects (pbi4_amd.fields) instead of numpy arrays; each
    converts to the reference's complex128 array on np.asarray().
  * Optional config keys: "precision" ("fp32" default, or "bf16x3" for the split-BF16 triangle
    GEMM), "maps_grid" (0 = band-limited grid chosen automatically, or an explicit size such as
    the full grid), "device" (GPU index).
  * IO.read_catalogues returns a DeviceCatalogue (np.asarray gives the reference's positions).
    Formats: the reference's gadget (binary), eos, minerva, flagship, quijote, molino and sancho
    readers (pbi4_amd.readers), plus "npz" (positions under 'pos', optional 'vel') and
    "synthetic" (seeded Zel'dovich catalogue, keys zel_n, zel_amp, zel_seed).
"""
import argparse
import json
import os
from json import JSONEncoder
from math import floor

import numpy as np
import torch

from . import _lib
from .core import Pbi4Error, default_context
from .fields import DeviceCatalogue, DeviceDensity, DeviceGrid, DeviceMultipole, half_from_array
from .readers import read_catalogue


class NumpyArrayEncoder(JSONEncoder):
    def default(self, obj):
        if isinstance(obj, np.ndarray):
            return obj.tolist()
        return JSONEncoder.default(self, obj)


def _ctx(params):
    return default_context(int(params.get("device", 0)))


def _precision(params):
    p = str(params.get("precision", "fp32")).lower()
    if p in ("fp32", "float32", "plain"):
        return _lib.PREC_FP32
    if p in ("bf16x3", "split-bf16", "bf16"):
        return _lib.PREC_BF16X3
    raise Pbi4Error(f"unknown precision {p!r}")


class IO:
    """Configuration, derived parameters and catalogue reading (pbi4/pbi4.py:22-245)."""

    def __init__(self, config_file=None, verbose=False, scriptmode=False):
        self.verbose = verbose
        if scriptmode:
            self.args = vars(self.parse_arguments())
        if config_file is None:
            self.config_file = self.args["config_file"]
        else:
            self.config_file = config_file
        with open(self.config_file) as f:
            self.params = json.load(f)
        if scriptmode:
            self.params.update({k: v for k, v in self.args.items() if v is not None})
        self.params.update(self.derived_params(self.params))
        # the reference rewrites its config file with the derived values (pbi4/pbi4.py:38-39)
        with open(self.params.get("config_file", self.config_file), "w") as f:
            json.dump(self.params, f, indent=4, cls=NumpyArrayEncoder)

    def parse_arguments(self):
        p = argparse.ArgumentParser()
        p.add_argument("--config_file", type=str, help="JSON configuration file")
        p.add_argument("--ifile", type=str, help="Input catalog file")
        p.add_argument("--dfile", type=str, help="Input/output density file")
        p.add_argument("--ps_outfile", type=str, help="Power spectrum output file")
        p.add_argument("--bs_outfile", type=str, help="Bispectrum output file")
        p.add_argument("--cat", type=str, help="Type of catalog")
        p.add_argument("--box", type=float, help="box size of the simulations in mpc/h")
        p.add_argument("--grid", type=float, help="size of the fft grid")
        p.add_argument("--interpolation", type=int, help="interpolation order 1-4")
        p.add_argument("--interlacing", action=argparse.BooleanOptionalAction)
        p.add_argument("--kbin", type=float, help="bin size in units of the fundamental frequency")
        p.add_argument("--kcenter", type=float, help="centre of the first bin in units of kF")
        p.add_argument("--iopen", action=argparse.BooleanOptionalAction)
        p.add_argument("--om", type=float)
        p.add_argument("--ol", type=float)
        p.add_argument("--z", type=float)
        p.add_argument("--multipoles", action=argparse.BooleanOptionalAction)
        p.add_argument("--los", nargs="+", type=float, help="Line Of Sight vector: --los X Y Z")
        p.add_argument("--cross", action=argparse.BooleanOptionalAction)
        p.add_argument("--cfile", type=str, help="Stem of the triangle counts file")
        p.add_argument("--npool", type=int, help="Number of threads (unused on the GPU)")
        p.add_argument("--savedensity", action=argparse.BooleanOptionalAction)
        p.add_argument("--precision", type=str, help="fp32 or bf16x3 triangle sums")
        p.add_argument("--maps_grid", type=int, help="grid of the bispectrum shell maps (0 = auto)")
        return p.parse_args()

    def read_params(self):
        if self.verbose:
            print("Input Parameters:")
            for key, value in self.params.items():
                print(key, " : ", value)
        return self.params

    def derived_params(self, params=None):
        """pbi4/pbi4.py:86-119."""
        if params is None:
            params = self.params
        params["los"] = np.array(params["los"])
        d = {"dk": params["kbin"] * 2.0 * np.pi / params["box"],
             "nbins_ps": floor((params["grid"] / 2 + params["kbin"] - params["kcenter"]) / params["kbin"]),
             "nbins_bisp": floor((params["grid"] / 3 + params["kbin"] - params["kcenter"]) / params["kbin"]),
             "kf": 2.0 * np.pi / params["box"],
             "kN": 2 * np.pi / params["box"] * params["grid"] / 2.0,
             "Hz": 100.0 * np.sqrt(params["om"] * (1.0 + params["z"]) ** 3 + params["ol"]),
             "los": params["los"] / (params["los"] @ params["los"])}
        kvec = np.array([(i - 1) * d["dk"] + params["kcenter"] * 2.0 * np.pi / params["box"]
                         for i in range(1, d["nbins_ps"] + 1)])
        ishift = int(params["kcenter"] / params["kbin"]) + (1 if params["iopen"] else 0)
        itot = 0
        for l in range(1, d["nbins_bisp"] + 1):
            for j in range(1, l + 1):
                imax = min(d["nbins_bisp"], j + l - 1 + ishift) if params.get("cross") else j
                for _ in range(max(1, l - j + 1 - ishift), imax + 1):
                    itot += 1
        d.update({"itot": itot, "kvec": kvec})
        return d

    def read_catalogues(self, params=None):
        """Normalised positions (3, npart) in [0, 1] as a DeviceCatalogue (pbi4/pbi4.py:121-245):
        the file is parsed on the host (pbi4_amd.readers), then the redshift-space shift along
        params['los'] (when params['multipoles'] and the format has one), the periodic wrap and
        the division by the box run on the device with the reference's arithmetic."""
        if params is None:
            params = self.params
        c = read_catalogue(params, self.verbose)
        npart = c.npart
        if self.verbose:
            print("Found %d %s." % (npart, c.kind))
            if c.kind != "particles":
                print("Number density: %.2e (h/Mpc)^3" % (npart / params["box"] ** 3))
        los = np.asarray(params["los"], dtype=np.float64)
        if c.rsd != _lib.RSD_NONE:
            print("Displacing along the LOS: (%.1f,%.1f,%.1f)" % (los[0], los[1], los[2]))
        ctx = _ctx(params)

        def dev(a):
            if isinstance(a, torch.Tensor):
                t = a.to(ctx.device)
            else:
                a = np.asarray(a)
                if a.dtype not in (np.float32, np.float64):
                    a = a.astype(np.float64)
                t = torch.from_numpy(np.ascontiguousarray(a)).to(ctx.device)
            return t.contiguous()

        pos = dev(c.pos)
        vel = dev(c.vel) if c.rsd != _lib.RSD_NONE else None
        out_dtype = pos.dtype if c.out_dtype is None else (
            torch.float64 if np.dtype(c.out_dtype) == np.float64 else torch.float32)
        norm = ctx.catalogue_positions(pos, vel, c.rsd, tuple(los), float(params.get("Hz", 1.0)),
                                       1.0 + float(params.get("z", 0.0)), float(params["box"]),
                                       out_dtype=out_dtype)
        print("Done.")
        return DeviceCatalogue(norm)


class DENSITIES:
    """Particle-to-grid assignment and the density field (pbi4/pbi4.py:248-396)."""

    def __init__(self, params, verbose=False):
        self.params = params
        self.verbose = verbose

    def read_density(self, params=None):
        """(npart, DeviceDensity) from the density file params['dfile'] written by save_density
        (pbi4/pbi4.py:257-288): header [grid, npart, kF, kN] and the complex128 (G, G, G) array.
        A file grid G above params['grid'] is reduced on the device to the lowest frequencies
        (powerI4.reduce_array); a smaller one is an error, as in the reference."""
        if params is None:
            params = self.params
        print("FILE: %s\n" % params["dfile"])
        print("Reading density...\n")
        ctx = _ctx(params)
        N = int(params["grid"])
        with np.load(params["dfile"], allow_pickle=False) as f:
            header = np.asarray(f["header"])
            print("Grid: %.0f" % header[0])
            print("Number of elements: %.0f\n" % header[1])
            G = int(header[0])
            if G < N:
                print("\nERROR: Density file has smaller grid size than requested. \n")
                raise SystemExit
            data = f["data"]
            if data.shape != (G, G, G):
                raise Pbi4Error(f"density data must be ({G}, {G}, {G}), got {data.shape}")
            # only the x <= N/2 slab is read downstream; [z][y][x] on the device
            slab = torch.from_numpy(np.ascontiguousarray(data[:N // 2 + 1].transpose(2, 1, 0),
                                                         dtype=np.complex128)).to(ctx.device)
        if G > N:
            print("Current grid size: %d" % G)
            print("Requested grid size: %d" % N)
            print("reducing density array to internal grid...")
        delta = ctx.reduce_array(slab, G, N, out_dtype=torch.complex64)
        npart = header[1]
        return npart, DeviceDensity(delta, N, int(params["interpolation"]), bool(params["interlacing"]),
                                    float(params["box"]), npart, upper_zero=G > N)

    def save_density(self, dcl, npart, params=None):
        """pbi4/pbi4.py:390-396: np.savez_compressed(dfile, header=[grid, npart, kF, kN],
        data=dcl) with dcl as the reference's complex128 (N, N, N) array."""
        if params is None:
            params = self.params
        print("FILE: %s" % params["dfile"])
        header = np.array([params["grid"], npart, params["kf"], params["kN"]])
        np.savez_compressed(params["dfile"], header=header, data=np.asarray(dcl))

    def assign_grid(self, pos, params=None, weight=None):
        """(npart, DeviceGrid): assign_double_grid / assign_single_grid (pbi4/pbi4.py:290-302)."""
        if params is None:
            params = self.params
        ctx = _ctx(params)
        npart = pos.shape[1]
        print("Assigning %d particles on the grid..." % npart)
        if weight is not None and not np.all(np.asarray(weight) == 1.0):
            raise Pbi4Error("per-particle weights other than 1 are not supported")
        if isinstance(pos, DeviceCatalogue):
            p = pos.pos.to(device=ctx.device, dtype=torch.float32).contiguous()
        elif isinstance(pos, torch.Tensor):
            p = pos.to(device=ctx.device, dtype=torch.float32).contiguous()
        else:
            p = torch.from_numpy(np.ascontiguousarray(pos, dtype=np.float32)).to(ctx.device)
        N = int(params["grid"])
        g = ctx.assign(p, N, int(params["interpolation"]), bool(params["interlacing"]))
        print("Done.")
        return npart, DeviceGrid(g, N, params["interpolation"], params["interlacing"], npart)

    def compute_density(self, npart, dcl, params=None):
        """DeviceDensity: dcl /= npart, FFT and fcomb (pbi4/pbi4.py:304-341)."""
        if params is None:
            params = self.params
        ctx = _ctx(params)
        N = int(params["grid"])
        print("Computing final density.")
        if self.verbose:
            print("Grid = %d" % N)
            print("Fundamental frequency = %f h/Mpc" % params["kf"])
            print("Nyquist frequency = %f h/Mpc" % params["kN"])
        inter = bool(params["interlacing"])
        if isinstance(dcl, DeviceGrid):
            grids = dcl.grids
        else:
            a = np.asarray(dcl)
            parts = [a.real, a.imag] if inter else [a.real]
            grids = torch.from_numpy(np.ascontiguousarray(
                np.stack([x.transpose(2, 1, 0) for x in parts]), dtype=np.float32)).to(ctx.device)
        delta = ctx.density(grids, N, int(params["interpolation"]), inter, float(params["box"]), npart)
        print("Done.\n")
        return DeviceDensity(delta, N, params["interpolation"], inter, params["box"], npart)

    def compute_density_multipoles(self, dcl, params=None):
        """(delta_2, delta_4) = (L_2(mu) delta, L_4(mu) delta) for the fixed line of sight
        (pbi4/pbi4.py:343-388)."""
        if params is None:
            params = self.params
        if not params["multipoles"]:
            print("WARNING: the parameter -multipoles- is set to False, so you should not compute multipoles!")
        if not isinstance(dcl, DeviceDensity):
            dcl = _density_from_array(dcl, params)
        print("Computing density multipoles...\n")
        return DeviceMultipole(dcl, 2, params["los"]), DeviceMultipole(dcl, 4, params["los"])


def _density_from_array(a, params):
    ctx = _ctx(params)
    N = int(params["grid"])
    return DeviceDensity(half_from_array(a, ctx.device), N, int(params["interpolation"]),
                         bool(params["interlacing"]), float(params["box"]), None)


def _as_density(x, params):
    return x if isinstance(x, DeviceDensity) else _density_from_array(x, params)


def _multipole_fields(dcl, dcl2, dcl4, params):
    """(delta2, delta4) device tensors for explicit fields, or (None, None) for L_l(mu)."""
    if dcl2 is None and dcl4 is None:
        return None, None
    if (isinstance(dcl2, DeviceMultipole) and isinstance(dcl4, DeviceMultipole)
            and dcl2.density is dcl and dcl4.density is dcl
            and np.allclose(dcl2.los, np.asarray(params["los"], dtype=np.float64))):
        return None, None
    dev = _ctx(params).device
    return half_from_array(dcl2, dev), half_from_array(dcl4, dev)


class MEASUREMENTS:
    """P(k) multipoles, triangle counts and bispectrum (pbi4/pbi4.py:398-594)."""

    def __init__(self, params, verbose=False):
        self.params = params
        self.verbose = verbose

    def powerspectrum(self, npart, dcl, dcl2=None, dcl4=None, params=None):
        if params is None:
            params = self.params
        ctx = _ctx(params)
        output_type = [("k", np.float32), ("avgk", np.float32), ("P0+PSN", np.float32),
                       ("P2", np.float32), ("P4", np.float32), ("Nmodes", np.float32),
                       ("PSN", np.float32)]
        output = np.zeros(len(params["kvec"]), dtype=output_type)
        psn = 1.0 / (npart * (2.0 * np.pi / params["box"]) ** 3)
        print("kmin = %f h/Mpc" % params["kvec"][0])
        print("kmax = %f h/Mpc" % params["kvec"][-1])
        print("Number of bins = %d" % params["nbins_ps"])
        N = int(params["grid"])
        d = _as_density(dcl, params)
        args = (d.delta, N, float(params["box"]), float(params["kbin"]), float(params["kcenter"]),
                int(params["nbins_ps"]))
        if params["multipoles"]:
            if params.get("cross"):
                print("ERROR: Cross power spectrum in redshift space not implemented yet!")
                raise SystemExit
            print("Computing redshift space power spectrum")
            f2, f4 = _multipole_fields(d, dcl2, dcl4, params)
            pout = ctx.pk(*args, True, tuple(params["los"]), delta2=f2, delta4=f4).cpu().numpy().T
            output["P0+PSN"], output["P2"], output["P4"], output["Nmodes"] = pout[1], pout[2], pout[3], pout[4]
        else:
            other = None
            if params.get("cross"):
                print("Computing reals space cross power spectrum.")
                if dcl2 is None:
                    print("ERROR: you did not include the second density field")
                    raise SystemExit
                other = _as_density(dcl2, params).delta
            else:
                print("Computing real space power spectrum")
            pout = ctx.pk(*args, False, delta_b=other).cpu().numpy().T
            output["P0+PSN"], output["Nmodes"] = pout[1], pout[2]
        output["k"] = params["kvec"]
        output["avgk"] = pout[0]
        output["PSN"] = psn
        print("Done.\n")
        return output

    def triangle_counts(self, params=None):
        """counts(itot) = N_tr N^3 (float64), exact; read from params['cfile'] if it exists
        (pbi4/pbi4.py:463-476)."""
        if params is None:
            params = self.params
        cfile = params.get("cfile")
        if cfile and os.path.exists(cfile) and os.path.getsize(cfile) > 0:
            print("Found triangle counts file.")
            print("FILE: %s" % cfile)
            with np.load(cfile, allow_pickle=False) as f:
                return np.asarray(f["counts"])
        if params.get("cross"):
            print("Computing triangle counts for cross bispectra")
        N = int(params["grid"])
        ntr = _ctx(params).count_triangles(int(params["nbins_bisp"]), N, float(params["kbin"]),
                                           float(params["kcenter"]), bool(params["iopen"]),
                                           bool(params.get("cross", False)), int(params["itot"]))
        return ntr.astype(np.float64) * float(N) ** 3

    def bispectrum(self, npart, counts, dcl, dcl2=None, dcl4=None, params=None):
        if params is None:
            params = self.params
        ctx = _ctx(params)
        N = int(params["grid"])
        d = _as_density(dcl, params)
        if params.get("cross"):
            # pbi4/pbi4.py:481-496: dcl2 is the second field; B uses it on the third bin
            if dcl2 is None:
                raise Pbi4Error("the cross bispectrum needs the second field as dcl2")
            d2 = _as_density(dcl2, params)
            desc = ctx.bk_desc(int(params["nbins_bisp"]), N, float(params["kbin"]),
                               float(params["kcenter"]), bool(params["iopen"]),
                               np.float32(params["kf"]), int(params["itot"]), int(npart), False,
                               tuple(params["los"]), _precision(params),
                               int(params.get("maps_grid", 0)))
            c = torch.from_numpy(np.ascontiguousarray(np.asarray(counts).astype(np.float32))).to(ctx.device)
            print("Computing cross bispectra")
            bout = ctx.bk_cross(d.delta, d2.delta, c, desc).cpu().numpy().T
            output = np.zeros(bout.shape[1], dtype=[(n, np.float32) for n in ("k1", "k2", "k3", "B0", "Ntr")])
            for r, n in enumerate(("k1", "k2", "k3", "B0", "Ntr")):
                output[n] = bout[r]
            print("Done.\n")
            return output
        mult = bool(params["multipoles"])
        desc = ctx.bk_desc(int(params["nbins_bisp"]), N, float(params["kbin"]),
                           float(params["kcenter"]), bool(params["iopen"]), np.float32(params["kf"]),
                           int(params["itot"]), int(npart), mult, tuple(params["los"]),
                           _precision(params), int(params.get("maps_grid", 0)))
        # f2py hands counts to the Fortran routine as REAL(4) (bispectrumI4.f90:8)
        c = torch.from_numpy(np.ascontiguousarray(np.asarray(counts).astype(np.float32))).to(ctx.device)
        names = ["k1", "k2", "k3", "Pk1", "Pk2", "Pk3", "B0+BSN", "BSN", "Ntr"]
        if mult:
            print("Computing redshift space bispectrum")
            names += ["B2", "B4"]
            f2, f4 = _multipole_fields(d, dcl2, dcl4, params)
        else:
            print("Computing real space bispectrum")
            f2 = f4 = None
        bout = ctx.bk(d.delta, c, desc, delta2=f2, delta4=f4).cpu().numpy().T
        output = np.zeros(bout.shape[1], dtype=[(n, np.float32) for n in names])
        for r, n in enumerate(names):
            output[n] = bout[r]
        print("Done.\n")
        return output

    def save_powerspectrum(self, pout, params=None):
        if params is None:
            params = self.params
        print("Saving power spectrum...\n")
        print("FILE: %s" % params["ps_outfile"])
        header = "k     avgk     P0+PSN      P2      P4     Nmodes    PSN\n"
        np.savetxt(params["ps_outfile"], np.array([pout["k"], pout["avgk"], pout["P0+PSN"], pout["P2"],
                                                   pout["P4"], pout["Nmodes"], pout["PSN"]]).T,
                   fmt=("%0.8e", "%0.8e", "%0.8e", "%0.8e", "%0.8e", "%d", "%0.8e"), header=header)
        print("Done.")

    def save_triangle_counts(self, counts, params=None):
        if params is None:
            params = self.params
        print("\nSaving triangle counts...")
        print("FILE %s" % params["cfile"])
        np.savez_compressed(params["cfile"], counts=counts)
        print("Done.\n")

    def save_bispectrum(self, bout, params=None):
        if params is None:
            params = self.params
        print("Saving bispectrum...\n")
        print("FILE: %s" % params["bs_outfile"])
        if params.get("cross"):
            header = "k1/kF k2/kF k3/kF      B0          N_tr\n"
            cols = ["k1", "k2", "k3", "B0", "Ntr"]
            fmt = ("%d", "%d", "%d", "%0.8e", "%0.8e")
        elif params["multipoles"] == 0:
            header = "k1/kF k2/kF k3/kF P(k1) P(k2)       P(k3)       B0+BSN        BSN    N_tr\n"
            cols = ["k1", "k2", "k3", "Pk1", "Pk2", "Pk3", "B0+BSN", "BSN", "Ntr"]
            fmt = ("%d", "%d", "%d") + ("%0.8e",) * 6
        else:
            header = ("k1/kF k2/kF k3/kF P(k1) P(k2)       P(k3)       B0+BSN        BSN    N_tr"
                      "    B2      B4\n")
            cols = ["k1", "k2", "k3", "Pk1", "Pk2", "Pk3", "B0+BSN", "BSN", "Ntr", "B2", "B4"]
            fmt = ("%d", "%d", "%d") + ("%0.8e",) * 8
        np.savetxt(params["bs_outfile"], np.array([bout[c] for c in cols]).T, fmt=fmt, header=header)
        print("Done.\n")
