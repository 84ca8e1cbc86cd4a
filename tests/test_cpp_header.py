"""The C++ processor header compiles on its own and its test program is built (CPU only)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_compiles_standalone(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text('#include "qmiv.hpp"\nint main() { qmiv::flt64V4 v{}; return (int)v[0]; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I" + os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_processor_test_program_built():
    assert os.path.exists(os.path.join(ROOT, "tests", "cpp", "_build", "test_processors"))
