import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def itcpd():
    """The engine package (libitcpd_hip.so must be built; there is no CPU fallback)."""
    import __graft_entry__ as G
    return G.load_package()


def lowrank(dims, R, noise, seed):
    """Sum of R rank-1 outer products of N(0,1) factors plus N(0, noise) (std) noise."""
    import numpy as np
    rng = np.random.default_rng(seed)
    true = [rng.standard_normal((d, R)) for d in dims]
    letters = "abcdefgh"[:len(dims)]
    T = np.einsum(",".join(l + "z" for l in letters) + "->" + letters, *true)
    return T + noise * rng.standard_normal(dims), true
