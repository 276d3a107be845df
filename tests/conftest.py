import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.join(ROOT, "tests")
for p in (ROOT, TESTS):  # the package and the shared test helpers (cli_golden)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle (our C17 restatement of engine/src/tpgp_rev.c); built on demand."""
    from oracle import bindings

    if not os.path.exists(bindings.ORACLE_SO):
        bindings.build()
    return bindings.CpuEngine("oracle")


@pytest.fixture(scope="session")
def reference():
    """The reference CPU engine compiled from its own sources, when it was built here."""
    from oracle import bindings

    if not os.path.exists(bindings.REF_SO):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    return bindings.CpuEngine("ref")


@pytest.fixture(scope="session")
def hip_ctx():
    from ddevol_amd.engine import Context

    ctx = Context(0)
    print(f"HIP device: {ctx.name}", flush=True)
    yield ctx
    ctx.close()
