"""The term compiler (host side of the ELL build): compiled masks and scalars
equal the operator-by-operator application bit for bit (tests/cpp/test_terms,
built by __graft_entry__.build())."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "test_terms")


def test_compiled_terms_equal_sequential_application():
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/test_terms is not built (run __graft_entry__.build())")
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "mismatches 0" in out.stdout
