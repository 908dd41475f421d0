import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def spp():
    import __graft_entry__ as ge
    return ge.load_package()


@pytest.fixture(scope="session")
def gpu_ctx(spp):
    ctx = spp.Context(0, "fp64")
    yield ctx
    ctx.close()


def to_product_state(spp, ostate):
    """Oracle state -> product FiniteState (same dense arrays and parities)."""
    L = ostate.lattice
    sites = [[spp.Site(ostate[(x, y)].a, ostate[(x, y)].p) for y in range(L.Ly)] for x in range(L.Lx)]
    return spp.FiniteState(spp.Lattice(L.Lx, L.Ly), sites, ostate.kind, beta=ostate.beta,
                           log_amplitude=ostate.log_amplitude)


def oracle_cfg(res, L):
    return np.array([[res["configuration"][x][y] for y in range(L.Ly)] for x in range(L.Lx)])
