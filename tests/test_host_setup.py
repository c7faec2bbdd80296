"""The product's host-side hierarchy construction (libcfdxhip, host-only
handle) against the oracle: aggregates, P, every Galerkin coarse matrix and
the certified damping are bit-identical."""
import numpy as np
import pytest

import cfdx_dev_amd as cx
from systems import box_mesh, pressure_system

CASES = [
    ("sa", (12, 10, 9), 0.15, 1.0, 0.0, dict()),
    ("sa", (20, 20, 20), 0.15, 1.0, 0.0, dict()),
    ("sa", (16, 24, 16), 0.10, 1.5, 0.1, dict(wall_layers=4, first_aspect=1000.0)),
    ("boomer", (12, 10, 9), 0.15, 1.0, 0.0, dict()),
    ("boomer", (16, 16, 16), 0.10, 1.5, 0.0, dict()),
]


@pytest.mark.parametrize("method,dims,jitter,sigma,filt,extra", CASES)
def test_hierarchy_bit_identical(oracle, method, dims, jitter, sigma, filt, extra):
    s = pressure_system(oracle, box_mesh(*dims, jitter=jitter, seed=3, **extra), sigma=sigma, seed=3)
    ao = oracle.amg(method, prolongator_filter=filt)
    assert ao.setup(s["n"], s["rowp"], s["cols"], s["vals"])
    ap = cx.NativeAMGPreconditioner(method=method, device=-1, prolongator_filter=filt, host_threads=4)
    assert ap.setup(s["n"], s["rowp"], s["cols"], s["vals"]), ap.last_error()
    assert ap.num_levels() == ao.num_levels()
    assert ap.smoother_omega() == ao.omega()
    for lvl in range(ao.num_levels()):
        for a, b in zip(ap.level_matrix(lvl), ao.level_matrix(lvl)):
            assert np.array_equal(a, b)
        if lvl + 1 < ao.num_levels():
            for a, b in zip(ap.level_transfer(lvl), ao.level_transfer(lvl)):
                assert np.array_equal(a, b)


def test_numeric_refresh_matches_oracle(oracle):
    s = pressure_system(oracle, box_mesh(10, 10, 10, jitter=0.15, seed=7), sigma=1.0, seed=7)
    ao = oracle.amg("sa")
    ap = cx.NativeAMGPreconditioner(device=-1)
    assert ao.setup(s["n"], s["rowp"], s["cols"], s["vals"]) and ap.setup(s["n"], s["rowp"], s["cols"], s["vals"])
    for scale in (2.0, 1.7, 0.5):
        assert ap.update_values(s["vals"] * scale) == ao.update_values(s["vals"] * scale)
    for lvl in range(ao.num_levels()):
        for a, b in zip(ap.level_matrix(lvl), ao.level_matrix(lvl)):
            assert np.array_equal(a, b)


def test_thread_count_does_not_change_results(oracle):
    s = pressure_system(oracle, box_mesh(14, 14, 14, jitter=0.15, seed=2), sigma=1.0, seed=2)
    outs = []
    for t in (1, 3, 8):
        ap = cx.NativeAMGPreconditioner(device=-1, host_threads=t)
        assert ap.setup(s["n"], s["rowp"], s["cols"], s["vals"])
        outs.append([ap.level_matrix(l) for l in range(ap.num_levels())])
    for o in outs[1:]:
        for A, B in zip(o, outs[0]):
            for a, b in zip(A, B):
                assert np.array_equal(a, b)
