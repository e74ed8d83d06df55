"""MatrixMarket I/O (core/src/Morpheus_MatrixMarket.hpp:62-209): the C++
drop-in's reader and writer run over the reference's fixture files
(tests/golden/mm_data, data files of core/tests/mm_data) with the cases of
core/tests/Test_MatrixMarket.hpp (tests/cpp/test_mmio.cpp; host only, no
GPU), and the Python reader gives the same containers."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from morpheus_amd import mmio
from morpheus_amd.capi import IOException, NotImplementedException

ROOT = Path(__file__).resolve().parent.parent
DATA = ROOT / "tests" / "golden" / "mm_data"
BIN = ROOT / "tests" / "cpp" / "test_mmio"


def test_cpp_reader_writer(tmp_path):
    r = subprocess.run(["make", "-s", "-C", str(ROOT / "tests" / "cpp"), str(BIN)],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([str(BIN), str(DATA), str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "0 failed tests" in r.stdout


def test_python_reader_symmetric_and_sorted():
    A = mmio.read_matrix_market(DATA / "symmetric_real.mtx")
    assert (A.nrows, A.ncols) == (4, 4)
    assert A.row_indices.tolist() == [0, 0, 1, 1, 2, 2, 3]
    assert A.column_indices.tolist() == [1, 2, 0, 2, 0, 1, 3]
    assert A.values.tolist() == [2.123, -45.345, 2.123, -4.8765, -45.345, -4.8765, 1.15]
    B = mmio.read_matrix_market(DATA / "unsymmetric_real_unsorted.mtx")
    assert B.row_indices.tolist() == [0, 0, 1, 1] and B.values.tolist() == [1.11, 2.22, 3.33, 4.44]
    P = mmio.read_matrix_market(DATA / "symmetric_pattern.mtx")
    assert P.values.tolist() == [1.0] * 7


def test_python_reader_arrays():
    M = mmio.read_matrix_market(DATA / "array_unsymmetric_real.mtx")
    assert M.tolist() == [[1.11, -2.22], [3.33, -4.44]]   # row-major, as the reference
    v = mmio.read_matrix_market(DATA / "vec_unsymmetric_integer.mtx")
    assert v.ravel().tolist() == [1, -2, 3, -4]


@pytest.mark.parametrize("name,exc", [
    ("complex.mtx", NotImplementedException), ("hermitian.mtx", NotImplementedException),
    ("skew-symmetric.mtx", NotImplementedException),
    ("mat_invalid_row_zero_index.mtx", IOException), ("mat_invalid_column_zero_index.mtx", IOException),
    ("mat_invalid_row_over_index.mtx", IOException), ("mat_invalid_column_over_index.mtx", IOException),
    ("mat_invalid_data_type.mtx", IOException), ("mat_unexpected_eof.mtx", IOException),
    ("mat_invalid_banner.mtx", IOException), ("mat_invalid_storage_format.mtx", IOException),
    ("mat_invalid_symmetry.mtx", IOException), ("mat_invalid_size.mtx", IOException),
    ("array_invalid_size.mtx", IOException), ("array_invalid_type_pattern.mtx", NotImplementedException),
    ("array_invalid_type_complex.mtx", NotImplementedException),
    ("array_invalid_type_custom.mtx", IOException), ("array_unexpected_eof.mtx", IOException),
    ("array_invalid_symmetry_symmetric.mtx", NotImplementedException)])
def test_python_reader_errors(name, exc):
    with pytest.raises(exc):
        mmio.read_matrix_market(DATA / name)


def test_python_write_roundtrip(tmp_path):
    from morpheus_amd import generators as gen
    from morpheus_amd.formats import csr_to_coo_host
    A = csr_to_coo_host(gen.powerlaw_rows(300, 200, 1.5, 6.0, 100, seed=2))
    p = tmp_path / "a.mtx"
    mmio.write_matrix_market(A, p, precision=17)
    B = mmio.read_matrix_market(p)
    assert np.array_equal(A.row_indices, B.row_indices)
    assert np.array_equal(A.column_indices, B.column_indices)
    assert np.array_equal(A.values, B.values)
