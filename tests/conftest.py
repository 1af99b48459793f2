import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def pytest_collection_finish(session):
    """Sessions that run GPU tests: torch's HIP runtime (the one bundled with
    the torch wheel) and the library's (the system ROCm) are two runtimes in
    one process.  torch's is initialised first, before the library has
    launched anything: a first torch initialisation that came after the
    library's GEMM, mixed and (T) tests (cooperative launches among them) has
    found no device."""
    if any(item.get_closest_marker("gpu") for item in session.items):
        try:
            import torch

            torch.cuda.is_available()
        except ImportError:
            pass


def load_molecule(name: str):
    """Fixture molecule -> (eps, B[P,p,q], nocc, e_scf)."""
    d = np.load(os.path.join(GOLDEN, name + ".npz"))
    eps = d["eps"]
    n = eps.shape[0]
    iu = np.triu_indices(n)
    Bpk = d["Bpk"]
    B = np.zeros((Bpk.shape[0], n, n))
    B[:, iu[0], iu[1]] = Bpk
    B[:, iu[1], iu[0]] = Bpk
    return eps, B, int(d["nocc"]), float(d["escf"])


def load_golden():
    import json

    with open(os.path.join(GOLDEN, "dalton_golden.json")) as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def golden():
    return load_golden()
