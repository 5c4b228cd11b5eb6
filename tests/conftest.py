import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_golden(name):
    """Load a golden fixture; chroma fixtures carry features, and C is rebuilt
    with the fixed-order FP64 cost (parflexdtw_amd.synth.cost_matrix_det)."""
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    if "C" not in d and "F1" in d:
        from parflexdtw_amd.synth import cost_matrix_det
        d["C"] = cost_matrix_det(d["F1"], d["F2"])
    return d


STAGE2_CASES = ["tiles_seed2", "parflex_seed3", "stage2_chroma300_L50",
                "stage2_chroma300_L64", "stage2_chroma420_L37", "stage2_chroma250_L400",
                "stage2_rand_L9", "stage2_rand_L5", "stage2_rand_4step", "stage2_nocand",
                "stage2_chroma600_L100", "stage2_chroma400_L5_cap"]


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.build()
    return oracle
