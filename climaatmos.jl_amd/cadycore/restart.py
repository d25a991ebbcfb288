"""Checkpoint and restart in ClimaCore's HDF5 format, through ``libcaio.so``.

Mirrors the reference's two ends of the path:

* ``save_state_to_disk(dycore, Y, t, output_dir)`` is
  ``save_state_to_disk_func(integrator, output_dir)``
  (src/callbacks/callbacks.jl:297-320): the file ``day{day}.{sec}.hdf5``
  holds Y under the field name ``"Y"`` and the root attributes ``time``
  (float32, the reference's FT) and ``atmos_model_hash``.
* ``get_state_restart(path, dycore, atmos_model_hash)`` is
  ``get_state_restart(restart_file, start_date, atmos_model_hash, comms_ctx)``
  (src/simulation/restart.jl:22-43): it returns ``(Y, t)`` and warns when the
  file records another model hash.

The arrays in the file are ClimaCore's parent arrays (VIJFH on the sphere, VF
on a column).  On the device they are converted to and from the native layout
with ``ca_unpack_state`` / ``ca_pack_state``.  The file functions
(``write_state`` / ``read_info`` / ``read_state``) work on host numpy arrays
and need no GPU.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import warnings

import numpy as np

from .build import IO_LIB

CA_IO_SPHERE, CA_IO_COLUMN = 0, 1
CA_IO_VIJFH, CA_IO_VIJHF, CA_IO_VF = 0, 1, 2
_LAYOUTS = {"VIJFH": CA_IO_VIJFH, "VIJHF": CA_IO_VIJHF, "VF": CA_IO_VF}
_LAYOUT_NAMES = {v: k for k, v in _LAYOUTS.items()}
EXPORTED = ("ca_io_write_state", "ca_io_read_info", "ca_io_read_state", "ca_io_last_error")


class CaIoGrid(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("layout", C.c_int32), ("nlev", C.c_int32), ("h_elem", C.c_int32),
        ("nq", C.c_int32), ("deep", C.c_int32), ("radius", C.c_double),
        ("z_face", C.POINTER(C.c_double)), ("elemorder", C.POINTER(C.c_int32)),
    ]


class CaIoInfo(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("layout", C.c_int32), ("nlev", C.c_int32), ("nq", C.c_int32),
        ("nelem", C.c_int32), ("h_elem", C.c_int32), ("float_bytes", C.c_int32), ("has_hash", C.c_int32),
        ("time", C.c_double), ("atmos_model_hash", C.c_uint64),
    ]


_lib = None


def load_io_library(path: str = IO_LIB):
    global _lib
    if _lib is not None and path == IO_LIB:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libcaio.so not built ({path}); run __graft_entry__.build() with libhdf5 present")
    lib = C.CDLL(path)
    fp = C.POINTER(C.c_float)
    lib.ca_io_write_state.argtypes = [C.c_char_p, C.POINTER(CaIoGrid), fp, fp, C.c_double, C.c_uint64]
    lib.ca_io_read_info.argtypes = [C.c_char_p, C.POINTER(CaIoInfo)]
    lib.ca_io_read_state.argtypes = [C.c_char_p, fp, C.c_int64, fp, C.c_int64]
    lib.ca_io_last_error.restype = C.c_char_p
    for nm in ("ca_io_write_state", "ca_io_read_info", "ca_io_read_state"):
        getattr(lib, nm).restype = C.c_int
    if path == IO_LIB:
        _lib = lib
    return lib


def _check(lib, st, what):
    if st != 0:
        raise RuntimeError(f"{what}: status {st}: {lib.ca_io_last_error().decode()}")


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def write_state(path, c, f, *, z_face, kind="sphere", h_elem=0, nq=4, radius=0.0, deep=True,
                elemorder=None, layout=None, time=0.0, atmos_model_hash=0):
    """Write host parent arrays: sphere ``c`` (Nh, 4, nq, nq, Nv) and ``f``
    (Nh, 1, nq, nq, Nv + 1) in C order (Julia's VIJFH); column ``c`` (4, Nv)
    and ``f`` (1, Nv + 1) (Julia's VF).  ``layout`` picks the file's data
    layout (sphere: "VIJFH" or "VIJHF")."""
    lib = load_io_library()
    z = np.ascontiguousarray(z_face, dtype=np.float64)
    nlev = z.size - 1
    c = np.ascontiguousarray(c, dtype=np.float32)
    f = np.ascontiguousarray(f, dtype=np.float32)
    g = CaIoGrid()
    if kind == "sphere":
        nh = 6 * h_elem * h_elem
        if c.shape != (nh, 4, nq, nq, nlev) or f.shape != (nh, 1, nq, nq, nlev + 1):
            raise ValueError(f"sphere arrays {c.shape}, {f.shape} do not match h_elem {h_elem}, nlev {nlev}")
        g.kind, g.layout = CA_IO_SPHERE, _LAYOUTS[layout or "VIJFH"]
    elif kind == "column":
        if c.shape != (4, nlev) or f.shape != (1, nlev + 1):
            raise ValueError(f"column arrays {c.shape}, {f.shape} do not match nlev {nlev}")
        g.kind, g.layout = CA_IO_COLUMN, _LAYOUTS[layout or "VF"]
    else:
        raise ValueError(f"unknown grid kind {kind!r}")
    g.nlev, g.h_elem, g.nq, g.deep, g.radius = nlev, h_elem, nq, int(bool(deep)), float(radius)
    g.z_face = z.ctypes.data_as(C.POINTER(C.c_double))
    eo = None
    if elemorder is not None:
        eo = np.ascontiguousarray(elemorder, dtype=np.int32)
        g.elemorder = eo.ctypes.data_as(C.POINTER(C.c_int32))
    st = lib.ca_io_write_state(os.fsencode(path), C.byref(g), _fptr(c), _fptr(f), float(time),
                               int(atmos_model_hash) & 0xFFFFFFFFFFFFFFFF)
    _check(lib, st, "ca_io_write_state")


def read_info(path) -> dict:
    lib = load_io_library()
    info = CaIoInfo()
    _check(lib, lib.ca_io_read_info(os.fsencode(path), C.byref(info)), "ca_io_read_info")
    d = {k: getattr(info, k) for k, _ in CaIoInfo._fields_}
    d["kind"] = "sphere" if info.kind == CA_IO_SPHERE else "column"
    d["layout"] = _LAYOUT_NAMES[info.layout]
    d["has_hash"] = bool(info.has_hash)
    return d


def read_state(path):
    """(c, f, info): the parent arrays of Y.c and Y.f as float32 host arrays
    in the shapes ``write_state`` takes (VIJHF files come back as VIJFH)."""
    lib = load_io_library()
    info = read_info(path)
    nl, nq, nh = info["nlev"], info["nq"], info["nelem"]
    if info["kind"] == "sphere":
        c = np.empty((nh, 4, nq, nq, nl), np.float32)
        f = np.empty((nh, 1, nq, nq, nl + 1), np.float32)
    else:
        c = np.empty((4, nl), np.float32)
        f = np.empty((1, nl + 1), np.float32)
    _check(lib, lib.ca_io_read_state(os.fsencode(path), _fptr(c), c.size, _fptr(f), f.size), "ca_io_read_state")
    return c, f, info


def checkpoint_name(t: float) -> str:
    """day{day}.{sec}.hdf5 from the elapsed time t [s] (callbacks.jl:304-307)."""
    t = float(np.float32(t))
    return f"day{math.floor(t / 86400.0)}.{math.floor(math.fmod(t, 86400.0))}.hdf5"


# --------------------------------------------------------- device state
def _vijfh_layouts(nlev: int, nh: int):
    """ca_layout x 5 of one buffer holding Y.c (VIJFH) followed by Y.f."""
    from .lib import CaLayout

    nc = 4 * 16 * nlev * nh
    lays = (CaLayout * 5)()
    for k in range(4):
        lays[k] = CaLayout(k * 16 * nlev, 1, nlev, 4 * nlev, 64 * nlev)
    lays[4] = CaLayout(nc, 1, nlev + 1, 4 * (nlev + 1), 16 * (nlev + 1))
    return lays, nc


def _grid_radius(dycore):
    g = dycore.grid
    return g.radius if g.radius > 0 else dycore.params.planet_radius


def save_state(dycore, Y, path, t, atmos_model_hash=0, column=0):
    """Write the native device state Y of ``dycore`` to ``path``.  A column
    handle writes its column ``column`` (the reference runs one column per
    simulation)."""
    torch = dycore.torch
    g = dycore.grid
    nl = g.nlev
    if g.kind == "sphere":
        nh = g.ncol // 16
        lays, nc = _vijfh_layouts(nl, nh)
        buf = torch.empty(nc + 16 * (nl + 1) * nh, dtype=torch.float32, device=dycore.device)
        dycore._check(dycore.lib.ca_unpack_state(dycore.h, dycore._ptr(Y), lays, C.c_void_p(buf.data_ptr()),
                                                 dycore._stream()), "ca_unpack_state")
        host = buf.cpu().numpy()
        c = host[:nc].reshape(nh, 4, 4, 4, nl)
        f = host[nc:].reshape(nh, 1, 4, 4, nl + 1)
        write_state(path, c, f, z_face=g.z_face, kind="sphere", h_elem=g.h_elem, nq=4,
                    radius=_grid_radius(dycore), deep=g.radius > 0, time=t, atmos_model_hash=atmos_model_hash)
        return
    if not 0 <= column < g.ncol:
        raise ValueError(f"column {column} outside [0, {g.ncol})")
    host = Y.cpu().numpy()
    nC = nl * g.ncol
    c = np.stack([host[k * nC:(k + 1) * nC].reshape(nl, g.ncol)[:, column] for k in range(4)])
    f = host[4 * nC:].reshape(nl + 1, g.ncol)[:, column][None, :]
    write_state(path, c, f, z_face=g.z_face, kind="column", time=t, atmos_model_hash=atmos_model_hash)


def save_state_to_disk(dycore, Y, t, output_dir, atmos_model_hash=0):
    """save_state_to_disk_func: writes ``output_dir/day{day}.{sec}.hdf5``."""
    path = os.path.join(output_dir, checkpoint_name(t))
    save_state(dycore, Y, path, t, atmos_model_hash)
    return path


def load_state(path, dycore, Y=None, column=None):
    """Read a checkpoint into a native device state of ``dycore`` (a new
    tensor, or ``Y`` in place).  A column file fills column ``column`` of Y,
    or every column when ``column`` is None.  Returns (Y, info)."""
    torch = dycore.torch
    g = dycore.grid
    c, f, info = read_state(path)
    if info["nlev"] != g.nlev:
        raise ValueError(f"{path}: {info['nlev']} levels, the handle has {g.nlev}")
    if Y is None:
        Y = dycore.zeros()
    if g.kind == "sphere":
        if info["kind"] != "sphere" or info["nelem"] * 16 != g.ncol or info["nq"] != 4:
            raise ValueError(f"{path}: {info['nelem']} elements of {info['nq']}^2 nodes do not fit the handle")
        lays, nc = _vijfh_layouts(g.nlev, g.ncol // 16)
        buf = torch.from_numpy(np.concatenate([c.reshape(-1), f.reshape(-1)])).to(dycore.device)
        dycore._check(dycore.lib.ca_pack_state(dycore.h, C.c_void_p(buf.data_ptr()), lays, dycore._ptr(Y),
                                               dycore._stream()), "ca_pack_state")
        torch.cuda.current_stream(dycore.device).synchronize()
        return Y, info
    if info["kind"] != "column":
        raise ValueError(f"{path}: a sphere checkpoint does not fit a column handle")
    nl, nC = g.nlev, g.nlev * g.ncol
    host = Y.cpu().numpy().copy()
    cols = slice(None) if column is None else slice(column, column + 1)
    for k in range(4):
        host[k * nC:(k + 1) * nC].reshape(nl, g.ncol)[:, cols] = c[k][:, None]
    host[4 * nC:].reshape(nl + 1, g.ncol)[:, cols] = f[0][:, None]
    Y.copy_(torch.from_numpy(host).to(dycore.device))
    return Y, info


def get_state_restart(path, dycore, atmos_model_hash=None):
    """(Y, t) from a checkpoint; warns when the file's model hash differs
    from ``atmos_model_hash`` (restart.jl:36-41)."""
    Y, info = load_state(path, dycore)
    if atmos_model_hash is not None and info["has_hash"] and info["atmos_model_hash"] != (
            int(atmos_model_hash) & 0xFFFFFFFFFFFFFFFF):
        warnings.warn(f"Restart file {path} was constructed with a different AtmosModel")
    return Y, info["time"]
