"""Host checks of device-code building blocks, built by __graft_entry__.build():
the term compiler (compiled masks and scalars equal the operator-by-operator
application bit for bit, tests/cpp/test_terms), the bit-sliced ring window
scan of the ring orbit minimum (equal to a plain scan of every rotation's
window, tests/cpp/test_ring_windows), the nibble-table colex rank of the ELL
kernels (equal to the rank's definition, tests/cpp/test_rank) and the host
tridiagonal solver of the Lanczos driver (closed-form constant tridiagonals and
a dense Jacobi solve, tests/cpp/test_tridiag)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(name):
    exe = os.path.join(ROOT, "tests", "cpp", name)
    if not os.path.exists(exe):
        pytest.fail(f"tests/cpp/{name} is not built (run __graft_entry__.build())")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "mismatches 0" in out.stdout


def test_compiled_terms_equal_sequential_application():
    run("test_terms")


def test_bit_sliced_ring_windows_equal_plain_scan():
    run("test_ring_windows")


def test_nibble_table_rank_equals_colex_rank():
    run("test_rank")


def test_tridiagonal_solver_equals_closed_forms_and_dense_solve():
    run("test_tridiag")
