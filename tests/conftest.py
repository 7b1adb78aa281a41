import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ctx2u1():
    from qcmaquis_amd import abi
    from qcmaquis_amd.engine import Context
    c = Context(0, abi.MQ_SYM_TWOU1)
    yield c
    c.close()


@pytest.fixture(scope="session")
def ctxsu2():
    from qcmaquis_amd import abi
    from qcmaquis_amd.engine import Context
    c = Context(0, abi.MQ_SYM_SU2U1)
    yield c
    c.close()
