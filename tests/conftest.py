import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import bind
    if not os.path.exists(bind.ORACLE_SO):
        import subprocess
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True, capture_output=True)
    return bind.oracle()


@pytest.fixture(scope="session")
def ref_lib():
    from oracle import bind
    r = bind.reference()
    if r is None:
        pytest.skip("oracle/_ref (compiled reference) not built here")
    return r


@pytest.fixture(scope="session")
def gpu():
    import dcns_amd
    ctx = dcns_amd.Context(0)
    yield ctx
    ctx.close()
