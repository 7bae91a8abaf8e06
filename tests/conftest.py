import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "bachelorarbeit-mlem_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def c1():
    """The 2D ring workload: 288 crystals, 64x64 voxels, seeded 5-ellipse
    phantom, Poisson counts summing to about 1e6."""
    import mlem_amd as M
    geo = M.RING2D_288()
    A = M.gen_system_matrix(geo, seed=1)
    x_true = M.gen_phantom(geo, seed=1, n_shapes=5, a_min=0.5, a_max=4.0)
    g = M.gen_sinogram(A, x_true, 1e6, seed=1)
    AT = M.transpose(A)
    return dict(geo=geo, A=A, AT=AT, g=g, x_true=x_true)
