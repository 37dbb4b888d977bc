"""Runs the C++ drop-in program (tests/cpp/dropin_test.cpp): the operator
classes of include/openfpm_dcpse used as the reference's own tests use theirs
(Dcpse_2D/3D_test, dcpse_op_tests, dcpse_op_save_load, dcpse_op_test_lap),
plus FP32 fields, in-place expressions, the row accessors, checkMomenta and FP64 applies on the sliced-ELL copy."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "dropin_test")


@pytest.mark.gpu
def test_cpp_dropin_program(cuda, tmp_path):
    assert os.path.exists(BIN), "tests/cpp/dropin_test is not built (run __graft_entry__.build())"
    proc = subprocess.run([BIN], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(proc.stdout)
    print(proc.stderr)
    assert proc.returncode == 0, proc.stdout[-4000:] + proc.stderr[-2000:]
    assert "0 failure(s)" in proc.stdout
    assert proc.stdout.count("[  OK  ]") == 10
