"""HDF5 tensor files: the reference's ``io/hdf5`` module and its raw k-space
loader, on the HDF5 C library through ctypes.

Follows ``cpp/corelib/lib/io/hdf5.cppm:14-19`` / ``hdf5.cpp:60-259``:

* ``import_tensor(path, dataset)`` -> CPU tensor;
* ``import_tensors(path, matchers=None)`` -> ``{"/group/name": tensor}`` for the
  datasets whose full path matches one of the regexes (``re.fullmatch``, as
  ``std::regex_match``), groups visited recursively in name order;
* ``export_tensor(tensor, path, dataset)`` overwrites the file with one dataset
  (``File::Overwrite``);
* ``export_tensors(mapping, path)`` writes several datasets into a new file
  (intermediate groups created), which the reference does one call at a time;
* ``import_raw_kspace(path)`` is the raw-data loader of
  ``cpp/corelib/interface/test_interface.cpp:45-112``: per encode e the
  coordinates ``/Kdata/K{X,Y,Z}_E{e}`` flattened and scaled by 3.141592 / 160
  (the reference's literal), the weights ``/Kdata/KW_E{e}`` and the coils
  ``/Kdata/KData_E{e}_C{c}``, c = 0, 1, ... while present, stacked to [C, M].

Element types, as in the reference: float32, float64, and complex64 /
complex128 stored as a compound of two float / double members (the real part
at the lower offset; the reference writes members "r", "i" and also reads
"real", "imag").  Anything else raises ``RuntimeError("disallowed dtype")``.

h5py and HighFive are not part of this environment; the HDF5 C library is
(``libhdf5``, found through ``ctypes.util.find_library``, ``HASTY_HDF5_LIB`` or
the conda prefix).  All tensors are host tensors: move them to the device
with ``.to(device)``.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import glob
import os
import re
import sys
from dataclasses import dataclass
from typing import Dict, List, Mapping, Optional, Sequence, Tuple

import numpy as np
import torch

_hid = ctypes.c_int64
_herr = ctypes.c_int
_hsize = ctypes.c_uint64

# enum values of the HDF5 1.10 C API
_H5F_ACC_RDONLY = 0
_H5F_ACC_TRUNC = 2
_H5P_DEFAULT = 0
_H5S_ALL = 0
_H5T_INTEGER, _H5T_FLOAT, _H5T_COMPOUND = 0, 1, 6
_H5I_GROUP, _H5I_DATASET = 2, 5
_H5_INDEX_NAME, _H5_ITER_INC = 0, 0


class _GInfo(ctypes.Structure):
    _fields_ = [("storage_type", ctypes.c_int), ("nlinks", _hsize), ("max_corder", ctypes.c_int64),
                ("mounted", ctypes.c_bool)]


_lib = None


def _find_library() -> str:
    cands = [os.environ.get("HASTY_HDF5_LIB"), ctypes.util.find_library("hdf5")]
    prefixes = [os.environ.get("CONDA_PREFIX"), "/opt/conda", sys.prefix]
    for p in prefixes:
        if p:
            cands += sorted(glob.glob(os.path.join(p, "lib", "libhdf5.so*")))
    for c in cands:
        if c and (os.path.exists(c) or not os.path.dirname(c)):
            try:
                ctypes.CDLL(c)
                return c
            except OSError:
                continue
    raise RuntimeError("the HDF5 C library (libhdf5) was not found; set HASTY_HDF5_LIB to its path")


def _h5():
    """The bound library (loaded and initialised once)."""
    global _lib
    if _lib is not None:
        return _lib
    lib = ctypes.CDLL(_find_library())
    sig = {
        "H5open": (_herr, []),
        "H5Eset_auto2": (_herr, [_hid, ctypes.c_void_p, ctypes.c_void_p]),
        "H5Fopen": (_hid, [ctypes.c_char_p, ctypes.c_uint, _hid]),
        "H5Fcreate": (_hid, [ctypes.c_char_p, ctypes.c_uint, _hid, _hid]),
        "H5Fclose": (_herr, [_hid]),
        "H5Dopen2": (_hid, [_hid, ctypes.c_char_p, _hid]),
        "H5Dcreate2": (_hid, [_hid, ctypes.c_char_p, _hid, _hid, _hid, _hid, _hid]),
        "H5Dget_type": (_hid, [_hid]),
        "H5Dget_space": (_hid, [_hid]),
        "H5Dread": (_herr, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
        "H5Dwrite": (_herr, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
        "H5Dclose": (_herr, [_hid]),
        "H5Gget_info": (_herr, [_hid, ctypes.POINTER(_GInfo)]),
        "H5Lget_name_by_idx": (ctypes.c_ssize_t, [_hid, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _hsize,
                                                  ctypes.c_char_p, ctypes.c_size_t, _hid]),
        "H5Oopen": (_hid, [_hid, ctypes.c_char_p, _hid]),
        "H5Oclose": (_herr, [_hid]),
        "H5Iget_type": (ctypes.c_int, [_hid]),
        "H5Pcreate": (_hid, [_hid]),
        "H5Pset_create_intermediate_group": (_herr, [_hid, ctypes.c_uint]),
        "H5Pclose": (_herr, [_hid]),
        "H5Screate_simple": (_hid, [ctypes.c_int, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)]),
        "H5Sget_simple_extent_ndims": (ctypes.c_int, [_hid]),
        "H5Sget_simple_extent_dims": (ctypes.c_int, [_hid, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)]),
        "H5Sclose": (_herr, [_hid]),
        "H5Tcreate": (_hid, [ctypes.c_int, ctypes.c_size_t]),
        "H5Tinsert": (_herr, [_hid, ctypes.c_char_p, ctypes.c_size_t, _hid]),
        "H5Tget_class": (ctypes.c_int, [_hid]),
        "H5Tget_size": (ctypes.c_size_t, [_hid]),
        "H5Tget_nmembers": (ctypes.c_int, [_hid]),
        "H5Tget_member_name": (ctypes.c_void_p, [_hid, ctypes.c_uint]),
        "H5Tget_member_offset": (ctypes.c_size_t, [_hid, ctypes.c_uint]),
        "H5Tget_member_type": (_hid, [_hid, ctypes.c_uint]),
        "H5Tclose": (_herr, [_hid]),
        "H5free_memory": (_herr, [ctypes.c_void_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    if lib.H5open() < 0:
        raise RuntimeError("H5open failed")
    lib.H5Eset_auto2(0, None, None)   # errors are reported here, not printed by the library
    g = lambda n: _hid.in_dll(lib, n).value   # noqa: E731
    lib.t_float, lib.t_double = g("H5T_NATIVE_FLOAT_g"), g("H5T_NATIVE_DOUBLE_g")
    lib.t_f32le, lib.t_f64le = g("H5T_IEEE_F32LE_g"), g("H5T_IEEE_F64LE_g")
    lib.t_int32 = g("H5T_NATIVE_INT_g")
    lib.p_lcpl = g("H5P_CLS_LINK_CREATE_ID_g")
    _lib = lib
    return lib


def _check(v, what):
    if v < 0:
        raise RuntimeError(f"HDF5: {what} failed")
    return v


def _member_name(lib, t, i) -> bytes:
    p = lib.H5Tget_member_name(t, i)
    try:
        return ctypes.string_at(p)
    finally:
        lib.H5free_memory(p)


def _read(lib, dset) -> torch.Tensor:
    ftype = _check(lib.H5Dget_type(dset), "H5Dget_type")
    space = _check(lib.H5Dget_space(dset), "H5Dget_space")
    mtype = None
    try:
        ndim = _check(lib.H5Sget_simple_extent_ndims(space), "H5Sget_simple_extent_ndims")
        dims = (_hsize * max(ndim, 1))()
        if ndim:
            _check(lib.H5Sget_simple_extent_dims(space, dims, None), "H5Sget_simple_extent_dims")
        shape = tuple(int(dims[i]) for i in range(ndim))
        cls, size = lib.H5Tget_class(ftype), lib.H5Tget_size(ftype)
        if cls == _H5T_FLOAT and size in (4, 8):
            dt, mem = (np.float32, lib.t_float) if size == 4 else (np.float64, lib.t_double)
        elif cls == _H5T_COMPOUND and size in (8, 16) and lib.H5Tget_nmembers(ftype) == 2:
            part = size // 2
            members = []
            for i in range(2):
                mt = lib.H5Tget_member_type(ftype, i)
                ok = lib.H5Tget_class(mt) == _H5T_FLOAT and lib.H5Tget_size(mt) == part
                lib.H5Tclose(mt)
                if not ok:
                    raise RuntimeError("disallowed dtype")
                members.append((lib.H5Tget_member_offset(ftype, i), _member_name(lib, ftype, i)))
            members.sort()
            if [m[0] for m in members] != [0, part]:
                raise RuntimeError("disallowed dtype")
            dt = np.complex64 if size == 8 else np.complex128
            mtype = _check(lib.H5Tcreate(_H5T_COMPOUND, size), "H5Tcreate")
            comp = lib.t_float if size == 8 else lib.t_double
            for off, name in members:
                _check(lib.H5Tinsert(mtype, name, off, comp), "H5Tinsert")
            mem = mtype
        else:
            raise RuntimeError("disallowed dtype")
        out = np.empty(shape, dtype=dt)
        if out.size:
            _check(lib.H5Dread(dset, mem, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, out.ctypes.data), "H5Dread")
        return torch.from_numpy(out)
    finally:
        if mtype is not None:
            lib.H5Tclose(mtype)
        lib.H5Sclose(space)
        lib.H5Tclose(ftype)


def _open(path: str):
    lib = _h5()
    if not os.path.exists(path):
        raise RuntimeError(f"HDF5 file {path} does not exist")
    return lib, _check(lib.H5Fopen(os.fsencode(path), _H5F_ACC_RDONLY, _H5P_DEFAULT), f"opening {path}")


def import_tensor(path: str, dataset: str) -> torch.Tensor:
    """The dataset ``dataset`` of the file at ``path`` (``hdf5.cpp:190-197``)."""
    lib, f = _open(path)
    try:
        d = lib.H5Dopen2(f, dataset.encode(), _H5P_DEFAULT)
        if d < 0:
            raise RuntimeError(f"HDF5: no dataset {dataset} in {path}")
        try:
            return _read(lib, d)
        finally:
            lib.H5Dclose(d)
    finally:
        lib.H5Fclose(f)


def import_tensors(path: str, matchers: Optional[Sequence] = None) -> Dict[str, torch.Tensor]:
    """Every dataset of the file whose full path ("/group/name") matches one of
    ``matchers`` (regexes, full match; all datasets when None), keyed by that
    path (``hdf5.cpp:200-234``)."""
    pats = None if matchers is None else [re.compile(m) if isinstance(m, str) else m for m in matchers]
    lib, f = _open(path)
    out: Dict[str, torch.Tensor] = {}

    def walk(gid, prefix):
        info = _GInfo()
        _check(lib.H5Gget_info(gid, ctypes.byref(info)), "H5Gget_info")
        for i in range(int(info.nlinks)):
            n = _check(lib.H5Lget_name_by_idx(gid, b".", _H5_INDEX_NAME, _H5_ITER_INC, i, None, 0, _H5P_DEFAULT),
                       "H5Lget_name_by_idx")
            buf = ctypes.create_string_buffer(n + 1)
            lib.H5Lget_name_by_idx(gid, b".", _H5_INDEX_NAME, _H5_ITER_INC, i, buf, n + 1, _H5P_DEFAULT)
            name = buf.value
            oid = _check(lib.H5Oopen(gid, name, _H5P_DEFAULT), "H5Oopen")
            try:
                kind = lib.H5Iget_type(oid)
                full = prefix + "/" + name.decode()
                if kind == _H5I_DATASET:
                    if pats is None or any(p.fullmatch(full) for p in pats):
                        out[full] = _read(lib, oid)
                elif kind == _H5I_GROUP:
                    walk(oid, full)
                else:
                    raise RuntimeError("unsupported object type")
            finally:
                lib.H5Oclose(oid)

    try:
        walk(f, "")
    finally:
        lib.H5Fclose(f)
    return out


def _write(lib, f, dataset: str, t: torch.Tensor, complex_names: Tuple[str, str]) -> None:
    a = np.require(t.detach().cpu().numpy(), requirements="C")   # keeps 0-d arrays 0-d
    ftype = mtype = None
    if a.dtype == np.float32:
        ftype, mtype = lib.t_f32le, lib.t_float
    elif a.dtype == np.float64:
        ftype, mtype = lib.t_f64le, lib.t_double
    elif a.dtype == np.int32:   # only for files that exercise the readers' type checks
        ftype = mtype = lib.t_int32
    elif a.dtype in (np.complex64, np.complex128):
        size = a.dtype.itemsize
        ftype = _check(lib.H5Tcreate(_H5T_COMPOUND, size), "H5Tcreate")
        mtype = _check(lib.H5Tcreate(_H5T_COMPOUND, size), "H5Tcreate")
        fp, mp = (lib.t_f32le, lib.t_float) if size == 8 else (lib.t_f64le, lib.t_double)
        for off, name in ((0, complex_names[0]), (size // 2, complex_names[1])):
            _check(lib.H5Tinsert(ftype, name.encode(), off, fp), "H5Tinsert")
            _check(lib.H5Tinsert(mtype, name.encode(), off, mp), "H5Tinsert")
    else:
        raise RuntimeError("disallowed dtype")
    dims = (_hsize * max(a.ndim, 1))(*a.shape)
    space = _check(lib.H5Screate_simple(a.ndim, dims, None), "H5Screate_simple")
    lcpl = _check(lib.H5Pcreate(lib.p_lcpl), "H5Pcreate")
    try:
        lib.H5Pset_create_intermediate_group(lcpl, 1)
        d = _check(lib.H5Dcreate2(f, dataset.encode(), ftype, space, lcpl, _H5P_DEFAULT, _H5P_DEFAULT),
                   f"creating dataset {dataset}")
        try:
            if a.size:
                _check(lib.H5Dwrite(d, mtype, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, a.ctypes.data), "H5Dwrite")
        finally:
            lib.H5Dclose(d)
    finally:
        lib.H5Pclose(lcpl)
        lib.H5Sclose(space)
        if a.dtype in (np.complex64, np.complex128):
            lib.H5Tclose(ftype)
            lib.H5Tclose(mtype)


def export_tensors(tensors: Mapping[str, torch.Tensor], path: str,
                   complex_names: Tuple[str, str] = ("r", "i")) -> None:
    """A new file at ``path`` holding every (dataset path -> tensor) of
    ``tensors``; groups on the way are created."""
    lib = _h5()
    f = _check(lib.H5Fcreate(os.fsencode(path), _H5F_ACC_TRUNC, _H5P_DEFAULT, _H5P_DEFAULT), f"creating {path}")
    try:
        for name, t in tensors.items():
            _write(lib, f, name, t, complex_names)
    finally:
        lib.H5Fclose(f)


def export_tensor(tensor: torch.Tensor, path: str, dataset: str) -> None:
    """Overwrite the file at ``path`` with the one dataset ``dataset``
    (``hdf5.cpp:237-258``; the tensor must be contiguous, as there)."""
    if not tensor.is_contiguous():
        raise RuntimeError("tensor must be contiguous")
    export_tensors({dataset: tensor}, path)


@dataclass
class RawEncode:
    """One encode of a raw k-space file: coordinates (kx, ky, kz) in radians
    in the reference's order, density weights [M], coil data [C, M]."""
    coords: List[torch.Tensor]
    weights: torch.Tensor
    kdata: torch.Tensor


RAW_COORD_SCALE = 3.141592 / 160.0   # the reference's literal (test_interface.cpp:72)


def import_raw_kspace(path: str, encodes: Optional[Sequence[int]] = None,
                      scale: float = RAW_COORD_SCALE) -> List[RawEncode]:
    """The raw-data loader of ``test_interface.cpp:36-112``: for every encode
    (all that have ``/Kdata/KX_E{e}`` when ``encodes`` is None) the scaled
    coordinates, the weights and the coil data stacked over ``_C0, _C1, ...``."""
    sets = import_tensors(path, [r"^/Kdata/KData_E.*", r"^/Kdata/KX_E.*", r"^/Kdata/KY_E.*",
                                 r"^/Kdata/KZ_E.*", r"^/Kdata/KW_E.*"])
    if encodes is None:
        encodes = sorted(int(m.group(1)) for k in sets for m in [re.fullmatch(r"/Kdata/KX_E(\d+)", k)] if m)
    out = []
    for e in encodes:
        try:
            coords = [(sets[f"/Kdata/K{a}_E{e}"].flatten() * scale) for a in "XYZ"]
            weights = sets[f"/Kdata/KW_E{e}"].flatten()
        except KeyError as err:
            raise RuntimeError(f"raw k-space file {path} has no dataset {err.args[0]}") from None
        coils = []
        while f"/Kdata/KData_E{e}_C{len(coils)}" in sets:
            coils.append(sets[f"/Kdata/KData_E{e}_C{len(coils)}"].flatten())
        if not coils:
            raise RuntimeError(f"raw k-space file {path} has no coil data for encode {e}")
        out.append(RawEncode(coords, weights, torch.stack(coils, 0)))
    return out
