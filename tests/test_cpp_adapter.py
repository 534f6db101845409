"""The C++ host adapter (fluidsim_amd/host/FluidGridHIP.hpp) over the C ABI."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "grid_api_test")
SRC = os.path.join(ROOT, "tests", "cpp", "grid_api_test.cpp")


def test_adapter_header_compiles_standalone(tmp_path):
    src = tmp_path / "inc.cpp"
    src.write_text('#include "fluidsim_amd/host/FluidGridHIP.hpp"\nint main() { return 0; }\n')
    subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-Werror", f"-I{ROOT}", str(src)], check=True)


def test_checkpoint_json_reader_and_writer(tmp_path):
    """fs_json.hpp on the host: bit-exact number round trips, nested values, grammar violations rejected."""
    exe = tmp_path / "json_test"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", f"-I{ROOT}",
                    os.path.join(ROOT, "tests", "cpp", "json_test.cpp"), "-o", str(exe)], check=True)
    proc = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(proc.stdout, proc.stderr)
    assert proc.returncode == 0, proc.stderr
    assert proc.stdout.count("PASS") == 3


@pytest.mark.gpu
def test_cpp_adapter_contract():
    assert os.path.exists(BIN), "build the C++ adapter test first (python -m fluidsim_amd.build)"
    proc = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(proc.stdout, proc.stderr)
    assert proc.returncode == 0, proc.stderr
    assert proc.stdout.count("PASS") == 9
