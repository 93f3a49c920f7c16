"""Shared fixtures: import paths for the product binding and the oracle, the
`gpu` marker, and the golden cases of tests/golden/golden.json."""
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "fherma-polymult_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

GOLDEN = os.path.join(ROOT, "tests", "golden", "golden.json")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def golden_cases():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


def golden_by_name():
    return {c["name"]: c for c in golden_cases()}


@pytest.fixture(scope="session")
def golden():
    return golden_by_name()
