import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    # A GPU call that never returns should name itself: dump every thread's
    # stack well before an outer watchdog would kill a silent run.
    import faulthandler

    faulthandler.dump_traceback_later(150, repeat=True)
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    # Build only what is missing: the oracle (gcc), the HIP library (hipcc)
    # and the command-line tools (g++).
    from tests import _oracle  # noqa: F401

    if not os.path.exists(_oracle.LIB_PATH):
        _oracle.build()
    from bbscite_amd import build

    if not os.path.exists(build.LIB):
        build.build_lib()
    if not (os.path.exists(build.CLI) and os.path.exists(build.QUALITY)):
        build.build_cli()
