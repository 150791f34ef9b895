import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")
ASTEX = os.path.join(GOLDEN, "astex")

# the step-0 energies of replica 0 of the 1a07C1 complex
# (reference doc/instructions.txt:172, src/common/util_print.cpp:351)
KAT_1A07C1 = np.array([0.6434, -0.0367, -0.2079, -0.1845, 0.8521, -0.8880, 0.0523, 0.1744, -1.0000, 0.7735])

# Parity tolerance of a normalised energy term against the FP64 oracle:
#   |x - y| <= REL_TOL * max(|y|, TERM_FLOOR)
# The floor keeps terms that pass through zero (electrostatics sums charges of
# both signs) from demanding more than FP32 can hold; 0.05 is about the
# smallest magnitude a term reaches in these workloads.
REL_TOL = 1e-4
TERM_FLOOR = 0.05


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def lib():
    from geauxdock_cs_v2_amd import lib as L
    return L.load()


@pytest.fixture(scope="session")
def kat_complex(lib):
    from geauxdock_cs_v2_amd import lib as L
    raw = L.load_job(os.path.join(ASTEX, "1a07C-1.pdb"), os.path.join(ASTEX, "1a07C1.sdf"),
                     os.path.join(ASTEX, "1a07C1-0.8.ff"), "1a07C1")
    return raw, raw.stage()


@pytest.fixture(scope="session")
def kat_complex11(lib):
    """1a07C with 11 protein conformations and 4 temperatures."""
    from geauxdock_cs_v2_amd import lib as L
    raw = L.load_job(os.path.join(ASTEX, "1a07C.pdb"), os.path.join(ASTEX, "1a07C1.sdf"),
                     os.path.join(ASTEX, "1a07C1-0.8.ff"), "1a07C1", n_tmp=4, temp_high=2.0, temp_low=0.3)
    return raw, raw.stage()


def rel_err(x, ref, floor=TERM_FLOOR):
    return np.abs(np.asarray(x, np.float64) - ref) / np.maximum(np.abs(ref), floor)
