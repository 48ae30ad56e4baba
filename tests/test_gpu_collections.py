"""C++ drop-in check: the Collections-API caller in tests/cpp/collection_parity.cpp runs
the gfx950 operators through Collection::ApplyOperator and compares with the oracle."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "tests", "cpp", "collection_parity")

pytestmark = pytest.mark.gpu


def test_collections_dropin_cpp():
    assert os.path.exists(BIN), "build tests/cpp first (make -C tests/cpp)"
    print("[case] tests/cpp/collection_parity", flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASS" in r.stdout
