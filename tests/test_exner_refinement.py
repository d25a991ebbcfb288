"""The float64 Exner function of the Newton stage (ca_dexner): a float32
estimate of q^(2/7) refined by one Newton step on y^7 = q^2, times
1 + (kappa - 2/7) ln q.  Restated in numpy and checked against q**kappa in
float64 over the pressure range of the workloads (q = p / p_ref from 1e-3 to
1.2), for the parameter set's kappa_d and for an exact 2/7."""
import numpy as np
import pytest

from cadycore import params as P


def dexner(q, kappa):
    qf = q.astype(np.float32)
    l2 = np.log2(qf)
    y0 = np.exp2(np.float32(2.0 / 7.0) * l2).astype(np.float64)
    y = (6.0 * y0 + q * q / y0 ** 6) / 7.0
    return y + y * ((kappa - 2.0 / 7.0) * (np.log(2.0) * l2.astype(np.float64)))


@pytest.mark.parametrize("kappa", [P.Params.from_toml().R_d / P.Params.from_toml().cp_d, 2.0 / 7.0])
def test_refined_exner_matches_float64_power(kappa):
    q = np.geomspace(1e-3, 1.2, 20001)
    exact = q ** kappa
    rel = np.abs(dexner(q, kappa) - exact) / exact
    assert rel.max() < 1e-12, rel.max()


def test_float32_estimate_alone_is_not_enough():
    # the refinement matters: the float32 estimate is off by ~1e-7
    q = np.geomspace(1e-3, 1.2, 2001)
    y0 = np.exp2(np.float32(2.0 / 7.0) * np.log2(q.astype(np.float32))).astype(np.float64)
    assert (np.abs(y0 - q ** (2.0 / 7.0)) / q ** (2.0 / 7.0)).max() > 1e-9
