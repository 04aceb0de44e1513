"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on CPU."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def lib():
    import gpupairhmm_amd as G
    if not os.path.exists(G.library_path()):
        G.build()
    return G.load_library()


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        O.build(with_ref=False)
    return O


@pytest.fixture(scope="session")
def engine(lib):
    import gpupairhmm_amd as G
    eng = G.Engine(mode=G.MODE_BATCHED)
    yield eng
    eng.close()


@pytest.fixture(scope="session")
def plain_engine(lib):
    import gpupairhmm_amd as G
    eng = G.Engine(mode=G.MODE_PLAIN)
    yield eng
    eng.close()


def say(msg: str) -> None:
    """Name the case before a GPU step (visible even if the step dies)."""
    print(msg, flush=True)
    sys.stdout.flush()
