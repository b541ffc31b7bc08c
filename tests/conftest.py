import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

EXAMPLES = os.path.join(ROOT, "tests", "golden", "examples")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libpdfp_hip)")


def example_path(name: str) -> str:
    return os.path.join(EXAMPLES, name + ".mps.gz")
