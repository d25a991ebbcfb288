"""Checkpoint and restart of the native state through ClimaCore's HDF5 format
(save_state_to_disk_func, src/callbacks/callbacks.jl:297-320; get_state_restart,
src/simulation/restart.jl:22-43), on the GPU: the device state goes through
ca_unpack_state to ClimaCore's VIJFH parent arrays, into the file, and back
through ca_pack_state.  A restart must continue bitwise as the uninterrupted
run does, and the file's arrays must be the ones the VIJFH descriptor of the
FieldVector hooks reads."""
import os
import warnings

import numpy as np
import pytest

from cadycore import build as B
from cadycore import geometry as G
from cadycore import initial as I
from cadycore import params as P

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(B.IO_LIB), reason="libcaio.so not built (no libhdf5)")]

PAR = P.Params.from_toml()


def _torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_sphere_restart_continues_bitwise(tmp_path):
    torch = _torch()
    from cadycore import restart as R
    from cadycore.lib import DryDycore

    g = G.cubed_sphere(2, G.tanh_stretched_faces(10, 30000.0, 300.0), PAR.planet_radius, PAR.Omega)
    num = P.Numerics(dt=100.0)
    d = DryDycore(g, PAR, num)
    try:
        Y = d.state(I.sphere_state(g, PAR, seed=2))
        d.dss(Y)
        d.step(Y, 100.0, 3)
        path = R.save_state_to_disk(d, Y, 300.0, str(tmp_path), atmos_model_hash=77)
        assert os.path.basename(path) == "day0.300.hdf5"
        info = R.read_info(path)
        assert (info["nelem"], info["nlev"], info["h_elem"], info["time"]) == (24, 10, 2, 300.0)
        # the file holds the VIJFH parent arrays of the hook boundary
        c, f, _ = R.read_state(path)
        cv, fv = d.to_vijfh(Y)
        assert np.array_equal(c, cv.cpu().numpy()) and np.array_equal(f, fv.cpu().numpy())
        Y2, t = R.get_state_restart(path, d, atmos_model_hash=77)
        torch.cuda.synchronize()
        assert t == 300.0
        assert torch.equal(Y2, Y)
        d.step(Y, 100.0, 2)
        d.step(Y2, 100.0, 2)
        torch.cuda.synchronize()
        assert torch.equal(Y2, Y), "restarted run differs from the uninterrupted one"
        with warnings.catch_warnings(record=True) as w:
            warnings.simplefilter("always")
            R.get_state_restart(path, d, atmos_model_hash=78)
        assert any("different AtmosModel" in str(x.message) for x in w)
    finally:
        d.close()


def test_column_checkpoint_round_trip(tmp_path):
    torch = _torch()
    from cadycore import restart as R
    from cadycore.lib import DryDycore

    g = G.columns(8, G.tanh_stretched_faces(20, 30000.0, 200.0))
    d = DryDycore(g, PAR, P.Numerics(dt=60.0))
    try:
        Y = d.state(I.column_state(g, PAR, seed=3))
        d.step(Y, 60.0, 2)
        p = str(tmp_path / "col.hdf5")
        R.save_state(d, Y, p, 120.0, column=5)
        info = R.read_info(p)
        assert (info["kind"], info["layout"], info["nlev"]) == ("column", "VF", 20)
        Yb = d.zeros()
        R.load_state(p, d, Y=Yb, column=5)
        torch.cuda.synchronize()
        nl, nC = g.nlev, g.nlev * g.ncol
        a, b = Y.cpu().numpy(), Yb.cpu().numpy()
        for k in range(4):
            assert np.array_equal(a[k * nC:(k + 1) * nC].reshape(nl, -1)[:, 5], b[k * nC:(k + 1) * nC].reshape(nl, -1)[:, 5])
        assert np.array_equal(a[4 * nC:].reshape(nl + 1, -1)[:, 5], b[4 * nC:].reshape(nl + 1, -1)[:, 5])
        assert not np.any(b[:nC].reshape(nl, -1)[:, 4]), "other columns untouched"
    finally:
        d.close()
