"""Host readers for the catalogue formats of IO.read_catalogues (pbi4/pbi4.py:121-245).

File parsing stays on the host; each reader returns a Catalogue: the positions in Mpc/h as an
(n, 3) array in the dtype the reference computes its shift and wrap in, the velocities when the
format and params ask for a redshift-space shift, the shift rule, and the dtype of the normalised
positions.  The shift, periodic wrap and normalisation then run on the device
(Context.catalogue_positions).

Formats (params['cat']), following the reference's reader branches:
  gadget    Gadget-2 snapshot, format 1 or 2 binary, either byte order, particle type 1
            (pbi4.py:143-154 with Utils/readgadget.py and Utils/readsnap.py).  As there, the file
            named by params['ifile'] (or ifile + '.0') is the one read; velocities carry the
            sqrt(a) factor when z != 0 (readsnap.py:424-425).  HDF5 snapshots need h5py, which
            is not installed here, and are rejected.
  eos       Rockstar halos, npz with 'mass', 'pos' (3, n), 'vel' (3, n) (pbi4.py:155-165).
  minerva   Minerva halo records <i4, 3f4 pos, 3f4 vel, f4 mass, i4> (pbi4.py:166-180).  The
            reference's redshift-space branch builds the velocity array with a call that numpy
            rejects; here the evident intent, s = x + v los, is applied.
  flagship  Euclid Flagship parquet columns halo_lm, x, y, z; real space only (pbi4.py:181-193).
  quijote   Quijote FoF halos, group_tab files of snapshot {z: 0->4, 0.5->3, 1->2, 2->1, 3->0},
            positions / 1e3 (pbi4.py:194-212 with Utils/readfof.py).  The reference keeps these
            in an object array of Python floats, so its shift and wrap run in float64.
  molino    Molino galaxies, text columns 0-2 (already in redshift space); redshift space only
            (pbi4.py:213-225).
  sancho    Sancho galaxies, npz with 'pos' (n, 3) (already in redshift space); redshift space
            only (pbi4.py:226-238).
  npz       positions under 'pos' as (n, 3) or (3, n), optional 'vel' (shifted like gadget).
  synthetic seeded Zel'dovich catalogue generated on the device (zel_n, zel_amp, zel_seed).
"""
import math
import os
import struct
from dataclasses import dataclass

import numpy as np

from ._lib import RSD_HZ, RSD_NONE, RSD_VEL, Pbi4Error


@dataclass
class Catalogue:
    pos: object                # (n, 3) positions, Mpc/h (numpy, or a device tensor for synthetic)
    vel: object = None         # (n, 3) velocities, or None
    rsd: int = RSD_NONE        # RSD_NONE, RSD_HZ (s = x + v los / H (1 + z)) or RSD_VEL (s = x + v los)
    out_dtype: object = None   # dtype of the normalised positions; None = pos's dtype
    kind: str = "particles"

    @property
    def npart(self):
        return int(self.pos.shape[0])


# ---------------------------------------------------------------------------- Gadget binary

_HEADER = struct.Struct("<6i6dddii6Iiidddd")      # the first 160 bytes of the 256-byte header


def gadget_filename(snapshot):
    """(filename, 'binary' | 'hdf5') as Utils/readgadget.py:7-18 resolves a snapshot name."""
    if os.path.exists(snapshot):
        return snapshot, ("hdf5" if snapshot[-4:] == "hdf5" else "binary")
    for suffix, kind in ((".0", "binary"), (".hdf5", "hdf5"), (".0.hdf5", "hdf5")):
        if os.path.exists(snapshot + suffix):
            return snapshot + suffix, kind
    raise Pbi4Error(f"snapshot {snapshot!r} not found")


@dataclass
class GadgetHeader:
    fmt: int
    endian: str
    npart: np.ndarray
    massarr: np.ndarray
    time: float
    redshift: float
    nall: np.ndarray
    filenum: int
    boxsize: float
    omega_m: float
    omega_l: float
    hubble: float


def gadget_header(filename):
    """Header of one Gadget binary file: format 2 files open with an 8-byte record holding the
    block name, format 1 files with the 256-byte header record (either byte order)."""
    with open(filename, "rb") as f:
        first = f.read(4)
        if len(first) < 4:
            raise Pbi4Error(f"{filename}: not a Gadget snapshot")
        for endian in ("<", ">"):
            (size,) = struct.unpack(endian + "i", first)
            if size in (8, 256):
                break
        else:
            raise Pbi4Error(f"{filename}: incorrect Gadget file format")
        fmt = 2 if size == 8 else 1
        if fmt == 2:
            f.seek(16, os.SEEK_CUR)          # name, next-block size, end marker, header marker
        raw = f.read(_HEADER.size)
    v = struct.unpack(endian + _HEADER.format[1:], raw)
    return GadgetHeader(fmt=fmt, endian=endian, npart=np.array(v[0:6], dtype=np.int64),
                        massarr=np.array(v[6:12]), time=v[12], redshift=v[13],
                        nall=np.array(v[16:22], dtype=np.int64), filenum=v[23], boxsize=v[24],
                        omega_m=v[25], omega_l=v[26], hubble=v[27])


def _find_block(filename, head, name, number):
    """(offset, size) of a data block: by its 4-character name in format 2, by its position
    (header = 1, POS = 2, VEL = 3) in format 1."""
    u32 = head.endian + "I"
    with open(filename, "rb") as f:
        f.seek(0, os.SEEK_END)
        end = f.tell()
        f.seek(0)
        k = 1
        while f.tell() < end:
            if head.fmt == 2:
                f.seek(4, os.SEEK_CUR)
                found = f.read(4).decode("latin-1") == name
                f.seek(8, os.SEEK_CUR)
            else:
                found = k == number
            (size,) = struct.unpack(u32, f.read(4))
            if found:
                return f.tell(), size
            f.seek(size, os.SEEK_CUR)
            (check,) = struct.unpack(u32, f.read(4))
            if check != size:
                raise Pbi4Error(f"{filename}: corrupt block record")
            k += 1
    raise Pbi4Error(f"{filename}: block {name!r} not found")


def read_gadget_block(filename, block, ptype=1):
    """POS or VEL of particle type ptype from one Gadget binary file as float32 (n, 3)
    (Utils/readsnap.py read_block on a single file)."""
    head = gadget_header(filename)
    number = {"POS ": 2, "VEL ": 3}[block]
    offset, size = _find_block(filename, head, block, number)
    total = int(head.npart.sum())
    if size != 12 * total:
        raise Pbi4Error(f"{filename}: block {block!r} holds {size} bytes, expected {12 * total}")
    skip = int(head.npart[:ptype].sum())
    n = int(head.npart[ptype])
    with open(filename, "rb") as f:
        f.seek(offset + 12 * skip)
        data = np.fromfile(f, dtype=np.dtype(head.endian + "f4"), count=3 * n)
    data = data.astype(np.float32).reshape(n, 3)
    if block == "VEL " and head.redshift != 0:
        data *= math.sqrt(head.time)          # float32 times a weak Python float, as numpy does
    return data


def read_gadget(ifile, ptype=1, velocities=False):
    filename, kind = gadget_filename(ifile)
    if kind == "hdf5":
        raise Pbi4Error("HDF5 Gadget snapshots need h5py, which is not installed")
    pos = read_gadget_block(filename, "POS ", ptype)
    vel = read_gadget_block(filename, "VEL ", ptype) if velocities else None
    return pos, vel


# ---------------------------------------------------------------------------- FoF halos

_QUIJOTE_SNAP = {0.0: 4, 0.5: 3, 1.0: 2, 2.0: 1, 3.0: 0}


def read_fof_groups(basedir, snapnum):
    """GroupMass, GroupPos, GroupVel (float32) from the group_tab files of a Gadget FoF
    catalogue: per file Ngroups, TotNgroups, Nids (int32), TotNids (uint64), Nfiles (uint32), then
    per group Len, Offset (int32), Mass, Pos[3], Vel[3], TLen[6], TMass[6] (float32) arrays."""
    ext = "%03d" % snapnum
    stem = os.path.join(basedir + "/groups_" + ext, "group_tab_" + ext + ".")
    mass, pos, vel = [], [], []
    fnb, nfiles = 0, 1
    while fnb < nfiles:
        with open(stem + str(fnb), "rb") as f:
            ng, _tot, _nids = np.fromfile(f, dtype=np.int32, count=3)
            np.fromfile(f, dtype=np.uint64, count=1)
            nfiles = int(np.fromfile(f, dtype=np.uint32, count=1)[0])
            if ng > 0:
                np.fromfile(f, dtype=np.int32, count=2 * ng)
                mass.append(np.fromfile(f, dtype=np.float32, count=ng))
                pos.append(np.fromfile(f, dtype=np.float32, count=3 * ng).reshape(ng, 3))
                vel.append(np.fromfile(f, dtype=np.float32, count=3 * ng).reshape(ng, 3))
                np.fromfile(f, dtype=np.float32, count=12 * ng)
            here = f.tell()
            f.seek(0, os.SEEK_END)
            if here != f.tell():
                raise Pbi4Error(f"{stem}{fnb}: finished reading before the end of the file")
        fnb += 1
    cat = lambda a, shape: np.concatenate(a) if a else np.zeros(shape, np.float32)  # noqa: E731
    return cat(mass, (0,)), cat(pos, (0, 3)), cat(vel, (0, 3))


# ---------------------------------------------------------------------------- dispatch


def _need(params, key):
    if key not in params or params[key] is None:
        raise Pbi4Error(f"params[{key!r}] is required for catalogue {params.get('cat')!r}")
    return params[key]


def read_catalogue(params, verbose=False):
    """Catalogue for params['cat'] with the reference's messages and SystemExit cases."""
    cat = params.get("cat")
    rsd_wanted = bool(params.get("multipoles"))
    names = {"gadget": "Gadget snapshot", "eos": "EOS Rockstar halos", "minerva": "Minerva halos",
             "flagship": "Euclid Flagship galaxies", "quijote": "Quijote halos",
             "molino": "Molino galaxies", "sancho": "Sancho galaxies", "npz": "npz catalogue"}
    if cat in names:
        print("Reading %s.\n" % names[cat])
        print("FILE: %s\n" % _need(params, "ifile"))
    if cat == "gadget":
        pos, vel = read_gadget(params["ifile"], 1, velocities=rsd_wanted)
        return Catalogue(pos, vel, RSD_HZ if rsd_wanted else RSD_NONE)
    if cat == "eos":
        with np.load(params["ifile"], allow_pickle=False) as f:
            mass, p = np.asarray(f["mass"]), np.asarray(f["pos"])
            v = np.asarray(f["vel"]) if rsd_wanted else None
        dt = np.result_type(mass, p)
        pos = np.ascontiguousarray(p.T, dtype=dt)
        return Catalogue(pos, None if v is None else np.ascontiguousarray(v.T),
                         RSD_HZ if rsd_wanted else RSD_NONE, kind="halos")
    if cat == "minerva":
        rec = np.dtype([("dummy", "i4"), ("pos", "f4", 3), ("Vel", "f4", 3), ("Mass", "f4"),
                        ("dummy2", "i4")])
        m = np.fromfile(params["ifile"], dtype=rec)
        pos = np.ascontiguousarray(m["pos"], dtype=np.float32)
        vel = np.ascontiguousarray(m["Vel"], dtype=np.float32) if rsd_wanted else None
        return Catalogue(pos, vel, RSD_VEL if rsd_wanted else RSD_NONE, kind="halos")
    if cat == "flagship":
        if rsd_wanted:
            print("redshift not implemented yet on %s" % cat)
            raise SystemExit
        import pandas as pd
        vec = np.asarray(pd.read_parquet(params["ifile"], engine="pyarrow",
                                         columns=["halo_lm", "x", "y", "z"]))
        return Catalogue(np.ascontiguousarray(vec[:, 1:4]), kind="halos")
    if cat == "quijote":
        z = params["z"]
        if z not in _QUIJOTE_SNAP:
            raise Pbi4Error(f"Quijote snapshots exist at z in {sorted(_QUIJOTE_SNAP)}, not {z}")
        snap = _QUIJOTE_SNAP[z]
        print("reading: ", params["ifile"], ", redshift: ", z, ", snapnum: ", snap)
        _mass, gpos, gvel = read_fof_groups(params["ifile"], snap)
        pos = gpos / np.float32(1e3)
        vel = gvel * np.float32(1.0 + z) if rsd_wanted else None
        return Catalogue(pos, vel, RSD_HZ if rsd_wanted else RSD_NONE, out_dtype=np.float64,
                         kind="halos")
    if cat in ("molino", "sancho"):
        if not rsd_wanted:
            print("real space not implemented yet on %s" % cat)
            raise SystemExit
        if cat == "molino":
            pos = np.ascontiguousarray(np.genfromtxt(params["ifile"])[:, [0, 1, 2, 13]][:, 0:3])
        else:
            with np.load(params["ifile"], allow_pickle=False) as f:
                pos = np.ascontiguousarray(np.asarray(f["pos"])[:, 0:3])
        return Catalogue(pos, kind="galaxies")
    if cat == "npz":
        with np.load(params["ifile"], allow_pickle=False) as f:
            pos = np.asarray(f["pos"])
            vel = np.asarray(f["vel"]) if "vel" in f.files else None
        if pos.shape[0] == 3 and pos.shape[1] != 3:
            pos = pos.T
            vel = vel.T if vel is not None else None
        use = rsd_wanted and vel is not None
        return Catalogue(np.ascontiguousarray(pos), np.ascontiguousarray(vel) if use else None,
                         RSD_HZ if use else RSD_NONE)
    if cat == "synthetic":
        import torch
        from .catalogue import zeldovich
        pos, vel = zeldovich(int(params.get("zel_n", 128)), box=float(params["box"]),
                             amp=float(params.get("zel_amp", 2.0e4)),
                             seed=int(params.get("zel_seed", 0)),
                             device="cuda" if torch.cuda.is_available() else "cpu")
        return Catalogue(pos.T.contiguous(), vel.T.contiguous() if rsd_wanted else None,
                         RSD_HZ if rsd_wanted else RSD_NONE)
    print("catalogue format %r is not supported (gadget, eos, minerva, flagship, quijote, "
          "molino, sancho, npz, synthetic)" % cat)
    raise SystemExit
