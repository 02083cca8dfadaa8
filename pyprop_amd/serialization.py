"""On-disk formats: wavefunctions and tensor potentials in HDF5.

Mirrors pyprop/serialization:
* SaveWavefunctionHDF / LoadWavefunctionHDF (WavefunctionHDF.py:2-117): the wavefunction
  is one complex128 dataset of the full shape at `datasetPath`, groups created on the way
  (CommonHDF.py:66-97), an existing node at that path replaced (RemoveExistingDataset,
  CommonHDF.py:15-33); loading checks the shape (LoadLocalSlab, CommonHDF.py:146-189).
* SaveTensorPotential / LoadTensorPotential (TensorPotentialHDF.py:10-127): the potential
  data at `groupPath/potential` and, per rank i, the global basis pairs at
  `groupPath/basisPairs<i>` with the attribute `storageId<i>` (SaveGeometryInfo, :160-168);
  loading refuses a file whose basis pairs differ from the potential's (CheckLocalSlab,
  :140-157) with the reference's message.  `groupPath` may not be "/".

The reference writes through PyTables (chunked CArrays); these files are written by
pyprop_amd.hdf5 with contiguous datasets and the same node names, dtypes and attributes,
and files of either kind are read.  The reference pickles its config object into the
`configObject` attribute; here `conf`, if given, is stored as text (the config file
contents), so that reading a file executes nothing.
"""
from __future__ import annotations

import os
from typing import Optional

import numpy as np

from .hdf5 import H5File, H5Writer, HDF5Error


def _norm_path(path: str) -> str:
    return path if path.startswith("/") else "/" + path


def _open_for_update(filename: str) -> H5Writer:
    return H5Writer.from_file(filename) if os.path.exists(filename) else H5Writer()


def _target(obj):
    """The array of a Wavefunction (its .data), or obj itself for an array or tensor
    (a numpy array's own .data is its buffer)."""
    return obj if isinstance(obj, np.ndarray) or hasattr(obj, "detach") else getattr(obj, "data", obj)


def _host_array(obj) -> np.ndarray:
    data = _target(obj)
    if hasattr(data, "detach"):  # a torch tensor (device or host)
        data = data.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(data))


def RemoveExistingDataset(filename: str, datasetPath: str) -> None:
    """Remove a node (and everything below it) if the file and the node exist
    (CommonHDF.py:15-33)."""
    if not os.path.exists(filename):
        return
    w = H5Writer.from_file(filename)
    w.remove(_norm_path(datasetPath))
    w.write(filename)


def SaveWavefunctionHDF(filename: str, datasetPath: str, psi, conf: Optional[str] = None) -> None:
    """Store psi (a Wavefunction, a torch tensor or an array) as complex128 at datasetPath
    (WavefunctionHDF.py:2-74)."""
    path = _norm_path(datasetPath)
    data = _host_array(psi).astype(np.complex128, copy=False)
    w = _open_for_update(filename)
    w.remove(path)
    attrs = {"CLASS": "CARRAY", "VERSION": "1.0", "TITLE": ""}
    if conf is not None:
        attrs["configObject"] = str(conf)
    w.create_dataset(path, data, attrs)
    w.write(filename)


def LoadWavefunctionHDF(filename: str, datasetPath: str, psi) -> None:
    """Load datasetPath into psi in place (WavefunctionHDF.py:77-100); the stored shape must
    equal psi's."""
    path = _norm_path(datasetPath)
    f = H5File(filename)
    if path not in f:
        raise HDF5Error("Dataset '%s' not found in file '%s'" % (path, filename))
    arr = f.read(path)
    target = _target(psi)
    if tuple(arr.shape) != tuple(target.shape):
        raise HDF5Error("Invalid shape on existing dataset. Got %s, expected %s" % (arr.shape, tuple(target.shape)))
    if hasattr(target, "copy_"):  # torch tensor, possibly on the device
        import torch
        target.copy_(torch.from_numpy(np.ascontiguousarray(arr, dtype=np.complex128)))
    else:
        target[...] = arr


def _geometries(potential):
    geoms = getattr(potential, "GeometryList", None) or getattr(potential, "geometries", None)
    if geoms is None:
        raise TypeError("potential has no geometry list")
    return geoms


def _potential_data(potential) -> np.ndarray:
    """Host complex128 data of a TensorPotential (regenerated on the device as it was set
    up) or of a TensorPotentialData."""
    if hasattr(potential, "device_potential"):
        return potential.device_potential().cpu().numpy()
    return np.asarray(potential.data, dtype=np.complex128)


def SaveTensorPotential(filename: str, groupPath: str, potential, conf: Optional[str] = None) -> None:
    """Store a tensor potential and its basis pairs under groupPath
    (TensorPotentialHDF.py:10-80, SaveGeometryInfo :160-168)."""
    if groupPath == "/":
        raise HDF5Error("Storing tensor potentials on '/' is not supported")
    group = _norm_path(groupPath).rstrip("/")
    data = _potential_data(potential)
    geoms = _geometries(potential)
    w = _open_for_update(filename)
    w.remove(group)
    w.create_group(group)
    attrs = {"CLASS": "CARRAY", "VERSION": "1.0", "TITLE": ""}
    w.create_dataset(group + "/potential", data, attrs)
    for i, g in enumerate(geoms):
        w.create_dataset(group + "/basisPairs%d" % i, np.ascontiguousarray(g.pairs, dtype=np.int32),
                         {"storageId%d" % i: g.storage_id})
    if conf is not None:
        w.set_attr(group, "configObject", str(conf))
    w.write(filename)


def CheckLocalSlab(filename: str, groupPath: str, geometryList) -> bool:
    """True when every rank's stored basis pairs equal the geometry's
    (TensorPotentialHDF.py:140-157)."""
    group = _norm_path(groupPath).rstrip("/")
    f = H5File(filename)
    for i, g in enumerate(geometryList):
        stored = f.read(group + "/basisPairs%d" % i)
        new = np.asarray(g.pairs)
        if stored.shape != new.shape or not np.array_equal(stored, new):
            return False
    return True


def LoadTensorPotential(filename: str, groupPath: str, potential) -> np.ndarray:
    """The stored potential data (host complex128) after checking its basis pairs
    (TensorPotentialHDF.py:83-127)."""
    if groupPath == "/":
        raise HDF5Error("Storing tensor potentials on '/' is not supported")
    group = _norm_path(groupPath).rstrip("/")
    geoms = _geometries(potential)
    if not CheckLocalSlab(filename, group, geoms):
        raise HDF5Error("Organization of basis pairs has changed from when this potential was generated. "
                        "Please regenerate potential")
    data = H5File(filename).read(group + "/potential")
    shape = tuple(g.pair_count for g in geoms)
    if tuple(data.shape) != shape:
        raise HDF5Error("Invalid shape on existing dataset. Got %s, expected %s" % (data.shape, shape))
    return data
