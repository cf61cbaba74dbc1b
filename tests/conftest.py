import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

# The oracle is an OpenMP library whose steps on small lattices are many short
# parallel regions.  Under pytest-xdist every worker would start one thread per
# CPU with spinning waits, and the workers then slow each other down a
# hundredfold; give each worker its share of the CPUs and passive waits.  This
# runs before the oracle library is loaded, when libgomp reads its settings.
if os.environ.get("PYTEST_XDIST_WORKER"):
    _workers = int(os.environ.get("PYTEST_XDIST_WORKER_COUNT", "1") or 1)
    os.environ["OMP_NUM_THREADS"] = str(max(1, (os.cpu_count() or 1) // max(1, _workers)))
    os.environ["OMP_WAIT_POLICY"] = "passive"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running check")


def pytest_report_header(config):
    return f"lbc_amd repo: {ROOT}"


@pytest.fixture(scope="session")
def ref_lib():
    import oracle
    r = oracle.reference()
    if r is None:
        pytest.skip("oracle/_ref/liblbm_ref.so not built (reference tree absent)")
    return r
