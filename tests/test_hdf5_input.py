"""HDF5 initial-condition and material input (the reference's VecLoad path,
src/tdyio.c:176-330) read from the reference's demo/richards/initial.h5, and
the checkpoint file (TDyIOOutputCheckpoint, src/tdyio.c:339-359)."""
import os

import numpy as np
import pytest

from tdycore_amd import hdf5io

HERE = os.path.dirname(os.path.abspath(__file__))
H5 = os.path.join(HERE, "golden", "initial.h5")


def test_lists_groups_and_datasets():
    assert hdf5io.list_members(H5) == ["IC", "Permeability", "PermeabilityX", "PermeabilityY", "PermeabilityZ",
                                       "Porosity", "fields"]
    assert hdf5io.list_members(H5, "fields") == ["IC"]


def test_reads_datasets_in_natural_cell_order():
    ic = hdf5io.read_dataset(H5, "IC")
    assert ic.shape == (60,) and ic.dtype == np.float64
    assert set(np.unique(ic)) == {99000.0, 99001.0}
    assert np.array_equal(hdf5io.read_dataset(H5, "fields/IC"), ic)
    np.testing.assert_array_equal(hdf5io.read_dataset(H5, "PermeabilityX"), np.full(60, 1.3e-9))
    np.testing.assert_array_equal(hdf5io.read_dataset(H5, "PermeabilityZ"), np.full(60, 1.1e-10))
    phi = hdf5io.read_dataset(H5, "Porosity")
    assert set(np.unique(phi)) <= {0.25, 0.3, 0.35, 0.38, 0.4}
    k = hdf5io.read_dataset(H5, "Permeability")
    assert set(np.unique(k)) == {1e-10, 1e-9}


def test_missing_member_and_foreign_file_raise(tmp_path):
    with pytest.raises(KeyError):
        hdf5io.read_dataset(H5, "fields/missing")
    with pytest.raises(ValueError):
        hdf5io.read_dataset(H5, "fields")
    p = tmp_path / "x.h5"
    p.write_bytes(b"not hdf5" * 8)
    with pytest.raises(ValueError):
        hdf5io.read_dataset(str(p), "IC")


def test_checkpoint_round_trip_and_layout(tmp_path):
    v = np.linspace(9.0e4, 1.0e5, 27)
    p = str(tmp_path / hdf5io.checkpoint_filename(3153.6))
    assert p.endswith("3.15360e+03_chk.h5")
    hdf5io.write_checkpoint(p, v, 42)
    assert hdf5io.list_members(p) == ["fields"] and hdf5io.list_members(p, "fields") == ["IC"]
    r = hdf5io.read_dataset(p, "fields/IC")
    assert r.dtype == np.float64 and np.array_equal(r, v)
    assert hdf5io.read_attribute(p, "time step") == 42
    # the dataset's header messages are the ones the HDF5 library wrote into initial.h5
    # (dataspace with max dims, IEEE little-endian double, fill value, contiguous layout)
    ref = hdf5io._File(open(H5, "rb").read())
    mine = hdf5io._File(open(p, "rb").read())
    kinds = lambda f, path: [t for t, _, _ in f.messages(f.lookup(path))]
    assert kinds(mine, "fields/IC") == [t for t in kinds(ref, "IC") if t in (1, 3, 5, 8)]
    dt = lambda f, path: [f.d[b:b + n] for t, b, n in f.messages(f.lookup(path)) if t == 3][0]
    assert dt(mine, "fields/IC") == dt(ref, "IC")
    with pytest.raises(KeyError):
        hdf5io.read_attribute(p, "missing")


H5DUMP = "/opt/conda/bin/h5dump"


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump (the HDF5 library's reader) is not installed")
def test_checkpoint_reads_back_with_the_hdf5_library(tmp_path):
    """The checkpoint as the HDF5 library sees it: h5dump lists the root
    attribute "time step" (int32 scalar) and "fields/IC" (float64 vector), and
    its binary dump of fields/IC equals the written values bit for bit."""
    import subprocess
    rng = np.random.default_rng(5)
    v = np.concatenate([rng.normal(1e5, 3e4, 997), [0.0, -0.0, 1.0 / 3.0]])
    p = str(tmp_path / "ck.h5")
    hdf5io.write_checkpoint(p, v, 42)
    head = subprocess.run([H5DUMP, "-H", p], check=True, capture_output=True, text=True).stdout
    assert 'ATTRIBUTE "time step"' in head and "H5T_STD_I32LE" in head and "SCALAR" in head
    assert 'GROUP "fields"' in head and 'DATASET "IC"' in head and "H5T_IEEE_F64LE" in head
    assert "( 1000 ) / ( 1000 )" in head
    att = subprocess.run([H5DUMP, "-a", "/time step", p], check=True, capture_output=True, text=True).stdout
    assert "(0): 42" in att
    out = str(tmp_path / "ic.bin")
    subprocess.run([H5DUMP, "-d", "/fields/IC", "-b", "LE", "-o", out, p], check=True, capture_output=True)
    back = np.fromfile(out, dtype="<f8")
    assert back.shape == v.shape and np.array_equal(back.view(np.uint64), v.view(np.uint64))
