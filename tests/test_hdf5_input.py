"""HDF5 initial-condition and material input (the reference's VecLoad path,
src/tdyio.c:176-330) read from the reference's demo/richards/initial.h5."""
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
