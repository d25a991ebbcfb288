"""Checkpoint files in ClimaCore's HDF5 format (libcaio.so), on the CPU.

The reference writes Y with ClimaCore's InputOutput.write!(HDF5Writer, Y, "Y")
plus the root attributes "time" and "atmos_model_hash"
(src/callbacks/callbacks.jl:297-320), and reads it back with
InputOutput.read_field and read_attribute (src/simulation/restart.jl:22-43).
ClimaCore 0.15.1 is not in the image and the reference holds no checkpoint
file, so the format is pinned by libhdf5's own tools instead:
- h5dump (the HDF5 library's reader, not libcaio's) lists the groups,
  datasets and attributes and the dataset values in file order, which must be
  Julia's column-major order of the parent arrays (HDF5 dimensions = Julia's
  reversed);
- a second writer compiled here from tests/h5_foreign_writer.c writes the
  variants a Float64 / VIJHF / HDF5.jl-variable-length-string file has, with
  values that name their own index, and libcaio must read them back.
What is not pinned: that ClimaCore's reader accepts every attribute spelling
(no ClimaCore here).
"""
import os
import re
import shutil
import subprocess
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "climaatmos.jl_amd"))
from cadycore import build as B  # noqa: E402

pytestmark = pytest.mark.skipif(not os.path.exists(B.IO_LIB), reason="libcaio.so not built (no libhdf5)")

H5DUMP = os.path.join(B.HDF5_ROOT, "bin", "h5dump")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _R():
    from cadycore import restart as R

    return R


def _sphere_arrays(h_elem, nlev, seed=0):
    rng = np.random.default_rng(seed)
    nh = 6 * h_elem * h_elem
    c = rng.normal(size=(nh, 4, 4, 4, nlev)).astype(np.float32)
    f = rng.normal(size=(nh, 1, 4, 4, nlev + 1)).astype(np.float32)
    return c, f


def _z(nlev):
    return np.linspace(0.0, 30000.0, nlev + 1) ** 1.0


def test_exports_every_declared_symbol():
    import ctypes
    import re

    hdr = open(os.path.join(ROOT, "include", "caio.h")).read()
    declared = set(re.findall(r"\b(ca_io_\w+)\s*\(", hdr))
    R = _R()
    assert declared == set(R.EXPORTED)
    lib = ctypes.CDLL(B.IO_LIB)
    for name in declared:
        assert hasattr(lib, name), name


@pytest.mark.parametrize("layout", ["VIJFH", "VIJHF"])
def test_sphere_round_trip(tmp_path, layout):
    R = _R()
    c, f = _sphere_arrays(2, 7)
    p = str(tmp_path / "s.hdf5")
    R.write_state(p, c, f, z_face=_z(7), kind="sphere", h_elem=2, radius=6.371229e6, layout=layout,
                  time=86400.0 * 3 + 17.25, atmos_model_hash=0xDEADBEEFCAFEF00D)
    info = R.read_info(p)
    assert info["kind"] == "sphere" and info["layout"] == layout
    assert (info["nlev"], info["nq"], info["nelem"], info["h_elem"]) == (7, 4, 24, 2)
    assert info["float_bytes"] == 4
    assert info["time"] == float(np.float32(86400.0 * 3 + 17.25))
    assert info["has_hash"] and info["atmos_model_hash"] == 0xDEADBEEFCAFEF00D
    c2, f2, _ = R.read_state(p)
    assert np.array_equal(c2, c) and np.array_equal(f2, f)


def test_column_round_trip(tmp_path):
    R = _R()
    rng = np.random.default_rng(1)
    c = rng.normal(size=(4, 10)).astype(np.float32)
    f = rng.normal(size=(1, 11)).astype(np.float32)
    p = str(tmp_path / "c.hdf5")
    R.write_state(p, c, f, z_face=_z(10), kind="column", time=60.0)
    info = R.read_info(p)
    assert (info["kind"], info["layout"], info["nlev"], info["nelem"]) == ("column", "VF", 10, 1)
    c2, f2, _ = R.read_state(p)
    assert np.array_equal(c2, c) and np.array_equal(f2, f)


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="no h5dump")
def test_h5dump_sees_climacore_structure(tmp_path):
    R = _R()
    c, f = _sphere_arrays(1, 5)
    p = str(tmp_path / "s.hdf5")
    R.write_state(p, c, f, z_face=_z(5), kind="sphere", h_elem=1, radius=6.371229e6, time=120.0,
                  atmos_model_hash=42)
    out = subprocess.run([H5DUMP, "-H", "-A", p], capture_output=True, text=True, check=True).stdout
    flat = " ".join(out.split())
    for s in ('GROUP "fields"', 'DATASET "Y.c"', 'DATASET "Y.f"', 'GROUP "Y"',
              'ATTRIBUTE "time" { DATATYPE H5T_IEEE_F32LE DATASPACE SCALAR',
              'ATTRIBUTE "atmos_model_hash" { DATATYPE H5T_STD_U64LE DATASPACE SCALAR',
              'DATASPACE SIMPLE { ( 6, 4, 4, 4, 5 ) / ( 6, 4, 4, 4, 5 ) }',
              'DATASPACE SIMPLE { ( 6, 1, 4, 4, 6 ) / ( 6, 1, 4, 4, 6 ) }',
              'GROUP "grids"', 'GROUP "extruded_finite_difference_grid"', 'GROUP "meshes"',
              'GROUP "cubedsphere"', 'GROUP "topologies"', 'GROUP "domains"'):
        assert s in flat, s
    # attribute values, read by the HDF5 library itself
    vals = subprocess.run([H5DUMP, "-a", "/fields/Y.c/data_layout", p], capture_output=True, text=True,
                          check=True).stdout
    assert '"VIJFH"' in vals
    vals = subprocess.run([H5DUMP, "-a", "/fields/Y/type", p], capture_output=True, text=True, check=True).stdout
    assert '"FieldVector"' in vals
    vals = subprocess.run([H5DUMP, "-a", "/atmos_model_hash", p], capture_output=True, text=True, check=True).stdout
    assert "42" in vals


def _h5dump_values(path, dset):
    out = subprocess.run([H5DUMP, "-d", dset, "-m", "%.9g", "-w", "0", "-y", path], capture_output=True,
                         text=True, check=True).stdout
    body = out[out.index("DATA {") + len("DATA {"):]
    body = body[:body.index("}")]
    return np.array([float(x) for x in body.replace(",", " ").split()])


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="no h5dump")
def test_h5dump_values_are_julia_column_major(tmp_path):
    """File order = Julia's column-major order of (Nv, Ni, Nj, F, Nh): level
    fastest, then i, j, component, element."""
    R = _R()
    nl, nh = 3, 6
    v, i, j, q, h = np.meshgrid(np.arange(nl), np.arange(4), np.arange(4), np.arange(4), np.arange(nh),
                                indexing="ij")
    julia = (1000 * q + 100 * h + 10 * j + i + v / 100.0).astype(np.float32)   # Julia dims (Nv, Ni, Nj, F, Nh)
    c = np.ascontiguousarray(julia.transpose(4, 3, 2, 1, 0))                  # our C-order (Nh, F, Nj, Ni, Nv)
    f = np.zeros((nh, 1, 4, 4, nl + 1), np.float32)
    p = str(tmp_path / "o.hdf5")
    R.write_state(p, c, f, z_face=_z(nl), kind="sphere", h_elem=1, radius=1.0)
    vals = _h5dump_values(p, "/fields/Y.c")
    assert np.array_equal(vals.astype(np.float32), julia.reshape(-1, order="F"))
    # the first values: (v, i, j, f, h) = (1..3, 1, 1, 1, 1), then i = 2
    assert np.allclose(vals[:4], [0.0, 0.01, 0.02, 1.0])


@pytest.fixture(scope="module")
def foreign_writer(tmp_path_factory):
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("no gcc")
    exe = str(tmp_path_factory.mktemp("h5w") / "h5_foreign_writer")
    inc, lib = os.path.join(B.HDF5_ROOT, "include"), os.path.join(B.HDF5_ROOT, "lib")
    subprocess.run([cc, "-O1", "-I", inc, os.path.join(ROOT, "tests", "h5_foreign_writer.c"), "-o", exe,
                    f"-L{lib}", f"-Wl,-rpath,{lib}", "-lhdf5"], check=True)
    return exe


def test_reads_foreign_float64_vijhf_file(tmp_path, foreign_writer):
    R = _R()
    p = str(tmp_path / "fx.hdf5")
    subprocess.run([foreign_writer, p, "1", "4"], check=True)
    info = R.read_info(p)
    assert info["layout"] == "VIJHF" and info["float_bytes"] == 8 and not info["has_hash"]
    assert info["time"] == 1234.5 and info["h_elem"] == 0 and info["nelem"] == 6
    c, f, _ = R.read_state(p)
    h, q, j, i, v = np.meshgrid(np.arange(6), np.arange(4), np.arange(4), np.arange(4), np.arange(4),
                                indexing="ij")
    assert np.array_equal(c, (1000.0 * q + 100 * h + 10 * j + i + v / 100.0).astype(np.float32))
    h, j, i, v = np.meshgrid(np.arange(6), np.arange(4), np.arange(4), np.arange(5), indexing="ij")
    assert np.array_equal(f[:, 0], (-(100.0 * h + 10 * j + i + v / 100.0)).astype(np.float32))


def test_rejects_moist_state_and_bad_sizes(tmp_path, foreign_writer):
    R = _R()
    p = str(tmp_path / "moist.hdf5")
    vt = "@NamedTuple{ρ::Float64, uₕ::ClimaCore.Geometry.Covariant12Vector{Float64}, ρe_tot::Float64, ρq_tot::Float64}"
    subprocess.run([foreign_writer, p, "1", "4", vt], check=True)
    with pytest.raises(RuntimeError, match="not the dry state"):
        R.read_info(p)
    with pytest.raises(RuntimeError, match="not an HDF5 file"):
        R.read_info(str(tmp_path / "missing.hdf5"))
    c, f = _sphere_arrays(1, 3)
    q = str(tmp_path / "ok.hdf5")
    R.write_state(q, c, f, z_face=_z(3), kind="sphere", h_elem=1, radius=1.0)
    lib = R.load_io_library()
    small = np.empty(10, np.float32)
    assert lib.ca_io_read_state(os.fsencode(q), R._fptr(small), small.size, R._fptr(small), small.size) == 1
    assert b"differ from the file" in lib.ca_io_last_error()
    with pytest.raises(RuntimeError, match="strictly increasing"):
        R.write_state(q, c, f, z_face=_z(3)[::-1], kind="sphere", h_elem=1, radius=1.0)


def test_old_namedtuple_spelling_is_read(tmp_path, foreign_writer):
    R = _R()
    p = str(tmp_path / "old.hdf5")
    vt = "NamedTuple{(:ρ, :uₕ, :ρe_tot), Tuple{Float64, ClimaCore.Geometry.Covariant12Vector{Float64}, Float64}}"
    subprocess.run([foreign_writer, p, "1", "2", vt], check=True)
    assert R.read_info(p)["nlev"] == 2


def test_elemorder_written_as_julia_matrix(tmp_path):
    if not os.path.exists(H5DUMP):
        pytest.skip("no h5dump")
    R = _R()
    c, f = _sphere_arrays(1, 2)
    eo = np.array([[1, 1, p] for p in (3, 1, 2, 6, 5, 4)], np.int32)   # (Nh, 3): x, y, panel
    p = str(tmp_path / "eo.hdf5")
    R.write_state(p, c, f, z_face=_z(2), kind="sphere", h_elem=1, radius=1.0, elemorder=eo)
    out = subprocess.run([H5DUMP, "-a", "/topologies/2d/elemorder", "-w", "0", p], capture_output=True, text=True,
                         check=True).stdout
    assert "( 6, 3 )" in out
    body = out[out.index("DATA {") + 6:]
    body = re.sub(r"\(\d+,\d+\):", " ", body[:body.index("}")])
    nums = [int(x) for x in body.replace(",", " ").split()]
    assert nums == eo.reshape(-1).tolist()
    with pytest.raises(RuntimeError, match="elemorder entry out of range"):
        R.write_state(p, c, f, z_face=_z(2), kind="sphere", h_elem=1, radius=1.0, elemorder=eo * 2)


def test_checkpoint_name():
    R = _R()
    assert R.checkpoint_name(0.0) == "day0.0.hdf5"
    assert R.checkpoint_name(2 * 86400.0 + 3600.5) == "day2.3600.hdf5"
