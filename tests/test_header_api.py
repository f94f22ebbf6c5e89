"""The drop-in C++ header API (include/cukd/): the reference's usage
patterns compile against it (compileKNN.cu / compileSpatialKNN.cu style),
and on the GPU device-side fcp/knn from a user kernel equal the batched
library calls (tests/cpp/header_api_test.cpp)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "header_api_test.bin")


def test_header_api_program_built():
    assert os.path.exists(BIN), "run __graft_entry__.build() (make -C tests/cpp)"


@pytest.mark.gpu
def test_header_api_program_runs():
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")
