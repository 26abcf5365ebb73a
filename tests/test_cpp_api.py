"""The C++17 host header (include/marbles_amd.hpp): it compiles on its own
with warnings as errors (no GPU needed), and on the GPU the test program
built from tests/cpp/test_cpp_api.cpp -- the reference's stage calls and
the fused step through the header, against the C oracle -- passes."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_compiles_standalone(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text('#include "marbles_amd.hpp"\nint main() { return marbles_amd::Stencil::N_MICRO_STATES == 27 ? 0 : 1; }\n')
    res = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                          "-I", os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


@pytest.mark.gpu
def test_cpp_program_matches_the_oracle():
    exe = os.path.join(ROOT, "tests", "cpp", "test_cpp_api")
    assert os.path.exists(exe), "build() compiles tests/cpp/test_cpp_api"
    res = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "cpp api ok" in res.stdout
