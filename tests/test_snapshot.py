"""The snapshot format (save_esdat, src/io/snap.F90:203-356, and the loaders
ohinit.F90:225-244, inipcl.F90:359-426, rescal.F90:186-192) on the CPU.

The layout is pinned by h5dump, a reader independent of this module: dataset
names, 32-bit integers and IEEE doubles (H5T_NATIVE_INTEGER / DOUBLE of
hdf5io.F90 wrt1i/wrt1d), and the dimension order the HDF5 Fortran interface
gives Fortran's sdoms(2,3) and rhobk(0:nx,0:ny,0:nz): (3, 2) and
(nz+1, ny+1, nx+1), indexed [axis][lower, upper] and [k][j][i]."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from ctca_amd import emg, snapshot

pytestmark = pytest.mark.skipif(not snapshot.hdf5_available(), reason="libhdf5 not found")

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump"
                                    if os.path.exists("/opt/conda/bin/h5dump") else None)


def records(counts, seed):
    rng = np.random.default_rng(seed)
    n = int(sum(counts))
    p = np.zeros(n, dtype=emg.PARTICLE)
    for f in ("x", "y", "z"):
        p[f] = rng.uniform(0, 8, n)
    for f in ("vx", "vy", "vz"):
        p[f] = rng.normal(size=n)
    p["spec"] = np.repeat(np.arange(1, len(counts) + 1), counts)
    return p


def h5dump(path, *args):
    return subprocess.run([H5DUMP, *args, path], capture_output=True, text=True, check=True).stdout


def test_round_trip_is_exact(tmp_path):
    counts = (700, 0, 300)
    p = records(counts, 1)
    p["x"][0] = np.nextafter(8.0, 0.0)
    p["vz"][1] = -0.0
    n = (8, 6, 4)
    rng = np.random.default_rng(2)
    rb = rng.normal(size=(n[2] + 1, n[1] + 1, n[0] + 1))
    chg = np.array([1.5, -2.25])
    path = snapshot.snapshot_file(str(tmp_path / snapshot.SAVE_DIR), 0)
    assert path.endswith(os.path.join("SNAPSHOT1", "esdat0000.h5"))
    snapshot.write_esdat(path, p, counts, n, rhobk=rb, chgacm=chg, rens=(0.5, 2.0, 3.0, 4.0), nnode=4)
    s = snapshot.read_esdat(path)
    assert (s.nnode, s.nspec, s.npc) == (4, 3, 2)
    assert s.np.tolist() == list(counts)
    assert s.sdoms.tolist() == [[0, 8], [0, 6], [0, 4]] and s.n == n
    assert s.rens.tolist() == [0.5, 2.0, 3.0, 4.0]
    for f in ("x", "y", "z", "vx", "vy", "vz", "preside", "spec"):
        got, ref = np.ascontiguousarray(s.pbuf[f]), np.ascontiguousarray(p[f])
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), f
    assert np.array_equal(s.rhobk, rb) and np.array_equal(s.chgacm, chg)


def test_chunked_columns(tmp_path, monkeypatch):
    """Columns longer than one transfer are written and read in pieces."""
    monkeypatch.setattr(snapshot, "_CHUNK", 64)
    counts = (150, 77)
    p = records(counts, 3)
    path = str(tmp_path / "esdat0000.h5")
    snapshot.write_esdat(path, p, counts, (8, 8, 8))
    s = snapshot.read_esdat(path)
    for f in ("x", "vy", "spec"):
        assert np.array_equal(s.pbuf[f], p[f]), f


def test_nnode_only_in_rank_zero(tmp_path):
    p = records((10,), 4)
    a, b = str(tmp_path / "esdat0000.h5"), str(tmp_path / "esdat0001.h5")
    snapshot.write_esdat(a, p, (10,), (4, 4, 4), nnode=2, myid=0)
    snapshot.write_esdat(b, p, (10,), (4, 4, 4), nnode=2, myid=1)
    assert snapshot.read_esdat(a).nnode == 2
    assert snapshot.read_esdat(b).nnode is None
    assert snapshot.snapshot_file("d", 12, snapshot.EMDAT) == os.path.join("d", "emdat0012.h5")


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not found")
def test_layout_read_by_h5dump(tmp_path):
    counts = (3, 2)
    p = records(counts, 5)
    n = (4, 3, 2)
    rb = np.zeros((n[2] + 1, n[1] + 1, n[0] + 1))
    rb[2, 1, 3] = 7.5           # node (i, j, k) = (3, 1, 2)
    path = str(tmp_path / "esdat0000.h5")
    snapshot.write_esdat(path, p, counts, n, rhobk=rb)
    head = h5dump(path, "-H")
    pat = r'DATASET "(\w+)" \{\s*DATATYPE\s+(\S+)\s*DATASPACE\s+SIMPLE \{ \( ([\d, ]+) \)'
    found = {name: (t, dims) for name, t, dims in re.findall(pat, head)}
    f64, i32 = "H5T_IEEE_F64LE", "H5T_STD_I32LE"
    want = {"nnode": (i32, "1"), "nspec": (i32, "1"), "np": (i32, "2"), "npc": (i32, "1"),
            "sdoms": (i32, "3, 2"), "rens": (f64, "4"), "px": (f64, "5"), "py": (f64, "5"),
            "pz": (f64, "5"), "pvx": (f64, "5"), "pvy": (f64, "5"), "pvz": (f64, "5"),
            "pres": (i32, "5"), "pis": (i32, "5"), "rhobk": (f64, "3, 4, 5")}
    for name, (t, dims) in want.items():
        assert found.get(name) == (t, dims), (name, found.get(name))
    assert 'DATASET "chgacm"' in head
    sd = h5dump(path, "-d", "sdoms")
    assert re.search(r"\(0,0\): 0, 4,\s*\(1,0\): 0, 3,\s*\(2,0\): 0, 2", sd)
    rbd = h5dump(path, "-d", "rhobk")
    assert re.search(r"\(2,1,0\): 0, 0, 0, 7.5, 0", rbd)
    pis = h5dump(path, "-d", "pis")
    assert re.search(r"\(0\): 1, 1, 1, 2, 2", pis)


def test_sdoms_mismatch_stops(tmp_path):
    path = str(tmp_path / "esdat0000.h5")
    snapshot.write_esdat(path, records((4,), 6), (4,), (8, 8, 8))
    s = snapshot.read_esdat(path)
    snapshot.check_sdoms(s, (8, 8, 8))
    with pytest.raises(snapshot.SnapshotError, match="sdoms is inconsistent"):
        snapshot.check_sdoms(s, (8, 8, 4))


def test_specsort_groups_by_species_stably():
    p = np.zeros(7, dtype=emg.PARTICLE)
    p["spec"] = [2, 1, 3, 1, 2, 1, 3]
    p["x"] = np.arange(7)
    q, counts = snapshot.specsort(p, 3)
    assert counts.tolist() == [3, 2, 2]
    assert q["spec"].tolist() == [1, 1, 1, 2, 2, 3, 3]
    assert q["x"].tolist() == [1, 3, 5, 0, 4, 2, 6]
    p["spec"][0] = 4
    with pytest.raises(snapshot.SnapshotError):
        snapshot.specsort(p, 3)


def test_rescale_to_the_new_units():
    """Positions scale by renr_new / renr_old and velocities by renv_new /
    renv_old (here 2 and 1/4, exact in binary); the file's own units give
    back the same values."""
    p = records((64,), 7)
    q = snapshot.rescale(p.copy(), (1.0, 2.0, 1.0, 1.0), renr=2.0, renv=0.5)
    assert np.array_equal(q["x"], 2.0 * p["x"]) and np.array_equal(q["vy"], 0.25 * p["vy"])
    same = snapshot.rescale(p.copy(), (0.5, 0.25, 1.0, 1.0), renr=0.5, renv=0.25)
    assert np.array_equal(same["z"], p["z"]) and np.array_equal(same["vx"], p["vx"])
    assert np.array_equal(snapshot.rescale_rhobk(np.ones(3), (1, 1, 4.0, 1), 2.0), np.full(3, 0.5))


def test_bad_inputs_are_refused(tmp_path):
    p = records((4,), 8)
    path = str(tmp_path / "esdat0000.h5")
    with pytest.raises(ValueError):
        snapshot.write_esdat(path, p, (5,), (4, 4, 4))
    with pytest.raises(ValueError):
        snapshot.write_esdat(path, p, (4,), (4, 4, 4), rhobk=np.zeros((2, 2, 2)))
    with pytest.raises(snapshot.SnapshotError):
        snapshot.read_esdat(str(tmp_path / "missing.h5"))
