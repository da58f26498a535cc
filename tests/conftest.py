import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "kokkos-particletransport_amd")
for p in (PKG, os.path.join(REPO, "oracle"), os.path.join(REPO, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def kpt_mod():
    import kpt

    kpt.load()
    return kpt


@pytest.fixture(scope="session")
def oracle_mod():
    import kpt_oracle

    kpt_oracle.load()
    return kpt_oracle
