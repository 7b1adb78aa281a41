"""The C++ forwarding layer include/qcmaquis_amd.hpp (contraction::Engine / solve<double> entry
points over the C ABI): the header compiles on its own and with the C ABI header; its test
program tests/cpp/engine_test (built by __graft_entry__.build) runs the engine's growth steps,
schedule and solve<double> on a small quantum-chemistry problem on the GPU (see that file)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "engine_test")


def test_header_compiles(tmp_path):
    src = tmp_path / "h.cpp"
    src.write_text('#include "qcmaquis_amd.hpp"\nint main() { qcmaquis_amd::DavidsonVector<double> v({3, 4});'
                   ' return int(v.n_blocks()) - 2; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_engine_program():
    assert os.path.exists(BIN), "tests/cpp/engine_test is not built (__graft_entry__.build)"
    print("running the C++ engine test", flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "ENGINE_TEST_OK" in r.stdout
