import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: longer CPU case")


def gpu_available() -> bool:
    try:
        from imexlbm_amd import capi
        return capi.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.build()
    return oracle


def golden(name: str) -> dict:
    """A fixture under tests/golden (data the reference's own tests hold)."""
    with open(os.path.join(ROOT, "tests", "golden", name)) as fh:
        return json.load(fh)


REGRESSION = golden("regression_deck.json")


def ke_matches_reference(ke_values) -> bool:
    """The first three reported KE values, to the digits the reference quotes."""
    want = REGRESSION["ke_first_reports"]
    return [float(f"{k:.4g}") for k in ke_values[:3]] == [float(f"{w:.4g}") for w in want]
