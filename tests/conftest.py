import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def masters():
    """Loader for the reference's master PNGs committed under tests/golden/masters
    (straight alpha, as pixie writes them), returned premultiplied like pixie's
    readImage does."""
    import numpy as np
    from PIL import Image as PILImage
    import oracle

    d = os.path.join(ROOT, "tests", "golden", "masters")

    def load(name, premultiply=True):
        a = np.asarray(PILImage.open(os.path.join(d, name.replace("/", "_"))).convert("RGBA")).copy()
        return oracle.to_premultiplied_alpha(a) if premultiply else a
    return load
