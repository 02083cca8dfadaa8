"""On-disk formats (pyprop/serialization): the HDF5 reader pinned on the reference's own
PyTables file, writer round trips, and the wavefunction / tensor-potential layouts."""
from pathlib import Path

import numpy as np
import pytest

from pyprop_amd.bspline import make_bspline
from pyprop_amd.hdf5 import H5File, H5Writer, HDF5Error
from pyprop_amd import serialization as ser

GOLDEN = Path(__file__).resolve().parent / "golden" / "helium_radial_groundstate.h5"


def test_reads_reference_pytables_file():
    """The reference's ground state (chunked CArray of the compound {r, i}) read here is
    S-normalised in the basis its config.ini names (43 B-splines of order 7, 40 linear
    breakpoints on [0, 25], 'zero' continuity) and symmetric under electron exchange:
    a wrong chunk placement, byte order, member offset or axis order breaks both."""
    f = H5File(str(GOLDEN))
    assert f.attrs("/")["CLASS"] == "GROUP"
    assert f.attrs("/wavefunction")["CLASS"] == "CARRAY"
    ds = f.node("/wavefunction")
    assert ds.layout[0] == "chunked"
    x = f.read("/wavefunction")
    assert x.dtype == np.complex128 and x.shape == (43, 43)
    b = make_bspline(7, "linear", 0.0, 25.0, 40, continuity="zero")
    S = b.overlap_band()
    assert abs(np.vdot(x, S @ x @ S.T) - 1.0) < 1e-12
    assert np.abs(x - x.T).max() < 1e-12


def test_rewrite_reference_file_identical(tmp_path):
    w = H5Writer.from_file(str(GOLDEN))
    out = tmp_path / "copy.h5"
    w.write(str(out))
    a, b = H5File(str(GOLDEN)), H5File(str(out))
    assert np.array_equal(a.read("/wavefunction"), b.read("/wavefunction"))
    assert a.attrs("/wavefunction") == b.attrs("/wavefunction")
    assert a.attrs("/") == b.attrs("/")
    assert b.node("/wavefunction").layout[0] == "contiguous"
    raw = out.read_bytes()
    assert raw[:8] == b"\x89HDF\r\n\x1a\n" and raw[8] == 0  # superblock version 0, as the reference's
    assert int.from_bytes(raw[40:48], "little") == len(raw)  # end-of-file address


def test_writer_round_trip_types_and_groups(tmp_path, rng):
    w = H5Writer()
    c = rng.standard_normal((3, 4, 5)) + 1j * rng.standard_normal((3, 4, 5))
    w.create_dataset("/a/b/c", c, {"CLASS": "CARRAY", "n": np.int64(7), "v": np.array([1.5, -2.0])})
    w.create_dataset("/f64", rng.standard_normal(11))
    w.create_dataset("/i32", np.arange(12, dtype=np.int32).reshape(6, 2), {"storageId0": "BandNH"})
    w.create_dataset("/c64", (rng.standard_normal(4) + 1j).astype(np.complex64))
    for k in range(30):  # more members than one node of the reference's K = 4 holds
        w.create_dataset("/many/x%02d" % k, np.full(2, float(k)))
    w.set_attr("/a", "title", "group a")
    p = tmp_path / "t.h5"
    w.write(str(p))
    f = H5File(str(p))
    assert np.array_equal(f.read("/a/b/c"), c)
    assert f.attrs("/a/b/c")["n"] == 7 and np.array_equal(f.attrs("/a/b/c")["v"], [1.5, -2.0])
    assert f.attrs("/a") == {"title": "group a"}
    assert f.read("/i32").dtype == np.int32 and f.attrs("/i32")["storageId0"] == "BandNH"
    assert f.read("/c64").dtype == np.complex64
    assert [f.read("/many/x%02d" % k)[0] for k in range(30)] == [float(k) for k in range(30)]
    with pytest.raises(KeyError):
        f.read("/missing")


def test_wavefunction_save_load(tmp_path, rng):
    p = str(tmp_path / "wf.h5")
    x = rng.standard_normal((4, 6, 6)) + 1j * rng.standard_normal((4, 6, 6))
    ser.SaveWavefunctionHDF(p, "/run/wavefunction", x, conf="[Propagation]\ntimestep = 0.01\n")
    ser.SaveWavefunctionHDF(p, "/run/wavefunction_2", 2 * x)
    ser.SaveWavefunctionHDF(p, "/run/wavefunction", 3 * x)  # replaces the node
    y = np.zeros_like(x)
    ser.LoadWavefunctionHDF(p, "/run/wavefunction", y)
    assert np.array_equal(y, 3 * x)
    ser.LoadWavefunctionHDF(p, "run/wavefunction_2", y)
    assert np.array_equal(y, 2 * x)
    with pytest.raises(HDF5Error, match="Invalid shape"):
        ser.LoadWavefunctionHDF(p, "/run/wavefunction", np.zeros((4, 6, 5), complex))
    with pytest.raises(HDF5Error, match="not found"):
        ser.LoadWavefunctionHDF(p, "/run/none", y)
    ser.RemoveExistingDataset(p, "/run/wavefunction_2")
    assert "/run/wavefunction_2" not in H5File(p) and "/run/wavefunction" in H5File(p)


def test_tensor_potential_save_load_and_pair_check(tmp_path):
    from pyprop_amd.systems import helium_system
    s = helium_system(lmax=2, L=(0, 1), order=5, n_breakpoints=10, rmax=12.0, n_inner=5, gamma=2.0)
    p = str(tmp_path / "pot.h5")
    for k, pot in enumerate(s.potentials):
        ser.SaveTensorPotential(p, "/potentials/%d" % k, pot)
    f = H5File(p)
    for k, pot in enumerate(s.potentials):
        for i, g in enumerate(pot.geometries):
            assert f.attrs("/potentials/%d/basisPairs%d" % (k, i)) == {"storageId%d" % i: g.storage_id}
        assert np.array_equal(ser.LoadTensorPotential(p, "/potentials/%d" % k, pot), pot.data)
    # a potential whose pairs differ from the stored ones is refused
    other = s.potentials[1]
    with pytest.raises(HDF5Error, match="Organization of basis pairs"):
        ser.LoadTensorPotential(p, "/potentials/0", other)
    with pytest.raises(HDF5Error):
        ser.SaveTensorPotential(p, "/", s.potentials[0])
