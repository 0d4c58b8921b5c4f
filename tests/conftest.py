import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# The test systems are small: without this, mixed-precision contexts would
# solve them on the FP64 path (products below 1e9 flops) and the precision
# ladder would go untested.  Read when a context is created.
os.environ.setdefault("TINCR_TDA_LADDER_MIN_FLOPS", "0")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: larger CPU cases")
