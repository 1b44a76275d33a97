"""Catalogue readers (pbi4_amd.readers) on files written here from the published layouts, the
reader dispatch of IO.read_catalogues (pbi4/pbi4.py:121-245) and the oracle of the device
front-end (shift, wrap, normalisation).  CPU only."""
import os
import struct

import numpy as np
import pytest

from oracle import pbi4_oracle as O
from pbi4_amd import readers as R
from pbi4_amd._lib import RSD_HZ, RSD_NONE, RSD_VEL


def _gadget_file(path, fmt, endian, npart, pos, vel, time=0.5, redshift=1.0):
    """A Gadget-2 snapshot: 256-byte header record, then POS, VEL and ID blocks for all types
    (format 2 puts an 8-byte name record before each block)."""
    head = struct.pack(endian + "6i6dddii6Iiidddd", *npart, *([0.0] * 6), time, redshift, 0, 0,
                       *npart, 0, 1, 1000.0, 0.3, 0.7, 0.7)
    head += b"\0" * (256 - len(head))
    blocks = [("HEAD", head), ("POS ", pos.astype(endian + "f4").tobytes()),
              ("VEL ", vel.astype(endian + "f4").tobytes()),
              ("ID  ", np.arange(len(pos), dtype=endian + "u4").tobytes())]
    with open(path, "wb") as f:
        for name, data in blocks:
            if fmt == 2:
                f.write(struct.pack(endian + "i", 8) + name.encode() +
                        struct.pack(endian + "i", len(data) + 8) + struct.pack(endian + "i", 8))
            f.write(struct.pack(endian + "I", len(data)) + data + struct.pack(endian + "I", len(data)))


@pytest.mark.parametrize("fmt,endian", [(1, "<"), (2, "<"), (1, ">"), (2, ">")])
def test_gadget_reader(tmp_path, fmt, endian):
    rng = np.random.default_rng(fmt * 10 + (endian == ">"))
    npart = [3, 5, 0, 2, 0, 0]
    pos = (rng.random((10, 3)) * 1000).astype(np.float32)
    vel = (rng.standard_normal((10, 3)) * 300).astype(np.float32)
    _gadget_file(str(tmp_path / "snap_003"), fmt, endian, npart, pos, vel, time=0.5, redshift=1.0)
    h = R.gadget_header(str(tmp_path / "snap_003"))
    assert h.fmt == fmt and list(h.npart) == npart and h.time == 0.5 and h.boxsize == 1000.0
    p, v = R.read_gadget(str(tmp_path / "snap_003"), 1, velocities=True)
    assert np.array_equal(p, pos[3:8])
    # velocities carry sqrt(a) when z != 0 (float32 product, as numpy computes it)
    assert np.array_equal(v, vel[3:8] * np.float32(np.sqrt(0.5)))
    _gadget_file(str(tmp_path / "snap_z0.0"), fmt, endian, npart, pos, vel, time=1.0, redshift=0.0)
    _, v0 = R.read_gadget(str(tmp_path / "snap_z0"), 1, velocities=True)   # resolved as name + '.0'
    assert np.array_equal(v0, vel[3:8])


def _fof_file(path, mass, pos, vel, tot, nfiles):
    ng = len(mass)
    with open(path, "wb") as f:
        np.array([ng, tot, 7 * ng], dtype=np.int32).tofile(f)
        np.array([7 * tot], dtype=np.uint64).tofile(f)
        np.array([nfiles], dtype=np.uint32).tofile(f)
        np.full(ng, 7, dtype=np.int32).tofile(f)
        np.arange(ng, dtype=np.int32).tofile(f)
        mass.astype(np.float32).tofile(f)
        pos.astype(np.float32).tofile(f)
        vel.astype(np.float32).tofile(f)
        np.zeros((ng, 12), dtype=np.float32).tofile(f)


def test_fof_reader_and_quijote_branch(tmp_path):
    rng = np.random.default_rng(2)
    mass = rng.random(9).astype(np.float32)
    pos = (rng.random((9, 3)) * 1e6).astype(np.float32)
    vel = (rng.standard_normal((9, 3)) * 300).astype(np.float32)
    d = tmp_path / "groups_003"
    d.mkdir()
    _fof_file(str(d / "group_tab_003.0"), mass[:4], pos[:4], vel[:4], 9, 3)
    _fof_file(str(d / "group_tab_003.1"), mass[:0], pos[:0], vel[:0], 9, 3)
    _fof_file(str(d / "group_tab_003.2"), mass[4:], pos[4:], vel[4:], 9, 3)
    m, p, v = R.read_fof_groups(str(tmp_path), 3)
    assert np.array_equal(m, mass) and np.array_equal(p, pos) and np.array_equal(v, vel)
    params = {"cat": "quijote", "ifile": str(tmp_path), "z": 0.5, "multipoles": True}
    c = R.read_catalogue(params)
    assert c.rsd == RSD_HZ and c.out_dtype == np.float64
    assert np.array_equal(c.pos, pos / np.float32(1e3)) and np.array_equal(c.vel, vel * np.float32(1.5))
    with pytest.raises(R.Pbi4Error):
        R.read_catalogue({**params, "z": 0.7})


def test_simple_formats(tmp_path):
    rng = np.random.default_rng(3)
    n = 20
    pos = rng.random((n, 3)) * 1000
    vel = rng.standard_normal((n, 3)) * 300
    # eos: npz with (3, n) rows; the positions take the common dtype of mass and pos
    np.savez(tmp_path / "eos.npz", mass=rng.random(n), pos=pos.T.astype(np.float32), vel=vel.T)
    c = R.read_catalogue({"cat": "eos", "ifile": str(tmp_path / "eos.npz"), "multipoles": True})
    assert c.pos.dtype == np.float64 and np.array_equal(c.pos, pos.astype(np.float32).astype(np.float64))
    assert c.rsd == RSD_HZ and np.array_equal(c.vel, vel)
    # minerva binary records
    rec = np.zeros(n, dtype=[("dummy", "i4"), ("pos", "f4", 3), ("Vel", "f4", 3), ("Mass", "f4"),
                             ("dummy2", "i4")])
    rec["pos"], rec["Vel"] = pos, vel
    rec.tofile(tmp_path / "minerva.bin")
    c = R.read_catalogue({"cat": "minerva", "ifile": str(tmp_path / "minerva.bin"), "multipoles": True})
    assert c.rsd == RSD_VEL and np.array_equal(c.pos, rec["pos"]) and np.array_equal(c.vel, rec["Vel"])
    # molino text (columns 0-2 of at least 14), sancho npz; both redshift space only
    np.savetxt(tmp_path / "molino.txt", np.hstack([pos, rng.random((n, 11))]))
    c = R.read_catalogue({"cat": "molino", "ifile": str(tmp_path / "molino.txt"), "multipoles": True})
    assert c.rsd == RSD_NONE and np.allclose(c.pos, pos, rtol=1e-15)
    np.savez(tmp_path / "sancho.npz", pos=pos.astype(np.float32))
    c = R.read_catalogue({"cat": "sancho", "ifile": str(tmp_path / "sancho.npz"), "multipoles": True})
    assert c.pos.dtype == np.float32 and np.array_equal(c.pos, pos.astype(np.float32))
    for cat, fname in (("molino", "molino.txt"), ("sancho", "sancho.npz")):
        with pytest.raises(SystemExit):
            R.read_catalogue({"cat": cat, "ifile": str(tmp_path / fname), "multipoles": False})
    # flagship parquet, real space only
    pd = pytest.importorskip("pandas")
    pytest.importorskip("pyarrow")
    pd.DataFrame({"halo_lm": rng.random(n) + 12, "x": pos[:, 0], "y": pos[:, 1], "z": pos[:, 2]}).to_parquet(
        tmp_path / "flag.parquet", engine="pyarrow")
    c = R.read_catalogue({"cat": "flagship", "ifile": str(tmp_path / "flag.parquet"), "multipoles": False})
    assert np.array_equal(c.pos, pos)
    with pytest.raises(SystemExit):
        R.read_catalogue({"cat": "flagship", "ifile": str(tmp_path / "flag.parquet"), "multipoles": True})
    with pytest.raises(SystemExit):
        R.read_catalogue({"cat": "unknown", "ifile": "x", "multipoles": False})


def test_oracle_catalogue_positions_invariants():
    # 0 <= x <= 1 and x * box is the shifted position modulo box, for both RSD rules and the
    # object-array branch
    rng = np.random.default_rng(4)
    box = 1000
    pos = ((rng.random((400, 3)) * 3 - 1) * box).astype(np.float32)
    vel = (rng.standard_normal((400, 3)) * 500).astype(np.float32)
    los = np.array([0.0, 0.6, 0.8])
    for rsd, obj in ((0, False), (1, False), (2, False), (1, True)):
        x = O.catalogue_positions(pos, vel, rsd, los, 130.0, 0.5, box, object_rows=obj)
        assert x.shape == (3, 400) and np.all((x >= 0) & (x <= 1))
        s = pos.astype(np.float64)
        if rsd == 1:
            s = s + vel * los / 130.0 * 1.5
        elif rsd == 2:
            s = s + vel * los
        assert np.allclose(x.T * box, np.mod(s, box), atol=2e-3)
