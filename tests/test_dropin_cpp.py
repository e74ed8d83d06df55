"""The C++ drop-in (morpheus_amd/include/Morpheus_AMD.hpp) tested the way
the reference tests its C++ API: tests/cpp/test_dropin.cpp mirrors
core/tests/Test_Multiply.hpp, core/tests/Serial/Test_Convert.hpp,
core/tests/Serial/Test_Convert_Dynamic.hpp, core/tests/Test_Copy.hpp and
core/tests/Test_DynamicMatrix.hpp, over the reference's 10x10 fixture and
oracle-checked random matrices.

CPU: the headers and the test program compile with the host compiler.
GPU: the built program runs and every case passes.
"""
from __future__ import annotations

import json
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CPP = ROOT / "tests" / "cpp"
BIN = CPP / "test_dropin"
GOLDEN = ROOT / "tests" / "golden" / "small_matrix.json"


def write_fixture_text(path: Path) -> None:
    """The JSON fixture as the line format test_dropin.cpp reads:
    ``C <NAME> <value>`` and ``<Fmt> <array> <i> [<j>] <value>``."""
    d = json.loads(GOLDEN.read_text())
    lines = [f"C {k} {int(v)}" for k, v in d["constants"].items()]
    for fmt, arrays in d["formats"].items():
        for key, entries in arrays.items():
            for idx, v in entries:
                lines.append(" ".join([fmt, key, *map(str, idx), repr(float(v))]))
    path.write_text("\n".join(lines) + "\n")


def test_fixture_text_roundtrip(tmp_path):
    p = tmp_path / "fx.txt"
    write_fixture_text(p)
    text = p.read_text().splitlines()
    assert "C SMALL_MATRIX_NNZ 32" in text
    # every value survives the text form exactly
    d = json.loads(GOLDEN.read_text())
    n_entries = sum(len(e) for a in d["formats"].values() for e in a.values())
    assert len(text) == len(d["constants"]) + n_entries
    for line in text:
        if not line.startswith("C "):
            assert float(line.split()[-1]) == float(repr(float(line.split()[-1])))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_headers_compile_with_host_compiler():
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
           "-Wno-unused-parameter", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           f"-I{ROOT / 'include'}", f"-I{ROOT / 'morpheus_amd' / 'include'}",
           f"-I{ROOT / 'oracle'}", str(CPP / "test_dropin.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]


@pytest.mark.gpu
def test_cpp_dropin_on_gpu(gpu, tmp_path):
    if not BIN.exists():
        pytest.fail(f"{BIN} is not built: run `make -C tests/cpp` (build() does)")
    fx = tmp_path / "fx.txt"
    write_fixture_text(fx)
    print("running", BIN, flush=True)
    r = subprocess.run([str(BIN), str(fx)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-6000:])
    print(r.stderr[-6000:])
    assert r.returncode == 0, "C++ drop-in tests failed"
    assert "0 failed tests" in r.stdout
