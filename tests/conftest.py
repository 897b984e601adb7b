import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# The AddressSanitizer runs (tests/test_asan.py) start pytest again with these
# set, so that the same tests load the instrumented builds instead.
if os.environ.get("GT_TEST_LIB"):
    from tracking_optimizations_amd import _lib as _gt_lib
    _gt_lib.LIB_PATH = os.environ["GT_TEST_LIB"]
if os.environ.get("GT_TEST_ORACLE_LIB"):
    from oracle import oracle as _gt_oracle
    _gt_oracle.use_library(os.environ["GT_TEST_ORACLE_LIB"])


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
