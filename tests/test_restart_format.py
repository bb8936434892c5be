"""VirtualFluids restart files (RestartObject.cpp) without a GPU.

* Known answers: the reference unit test's data {{1,2,3},{4,5,6}}
  (RestartObjectTests.cpp, saveAndLoad) gives, in the binary form, the int32
  level count, then per level an int32 count and the float32 values; in the
  text form one line per level, each value followed by a space.
* The text form writes std::ostream's default six significant digits.
* NaN is written as 0 (RestartObject::serialize).
* The Python writer and the C++ host layer's RestartObject classes write the
  same bytes, and each reads what the other wrote.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from tests.test_host_abi import build_host_test
from virtualfluids_amd import restart


def test_binary_known_answer(tmp_path):
    name = tmp_path / "cp"
    restart.write(name, [[1, 2, 3], [4, 5, 6]], "binary")
    expected = struct.pack("<i", 2) + struct.pack("<i", 3) + struct.pack("<3f", 1, 2, 3) + \
        struct.pack("<i", 3) + struct.pack("<3f", 4, 5, 6)
    assert (tmp_path / "cp.bin").read_bytes() == expected
    back = restart.read(name, "binary")
    assert [a.tolist() for a in back] == [[1, 2, 3], [4, 5, 6]]


def test_binary_double(tmp_path):
    name = tmp_path / "cp"
    data = [np.array([0.1, -2.5e-300, 1e300]), np.array([], dtype=np.float64)]
    restart.write(name, data, "binary", real=np.float64)
    raw = (tmp_path / "cp.bin").read_bytes()
    assert raw == struct.pack("<ii3d", 2, 3, 0.1, -2.5e-300, 1e300) + struct.pack("<i", 0)
    back = restart.read(name, "binary", real=np.float64)
    assert np.array_equal(back[0], data[0]) and back[1].size == 0


def test_ascii_known_answer_and_format(tmp_path):
    name = tmp_path / "cp"
    restart.write(name, [[1, 2, 3], [4, 5, 6], [0.123456789, -1e-7, 12345678.0, -0.0]], "ascii")
    text = (tmp_path / "cp.txt").read_text()
    assert text == "1 2 3 \n4 5 6 \n0.123457 -1e-07 1.23457e+07 -0 \n"
    back = restart.read(name, "ascii")
    assert back[2].dtype == np.float32
    assert np.allclose(back[2], np.float32([0.123457, -1e-07, 1.23457e+07, 0.0]), rtol=0, atol=0)


def test_nan_written_as_zero(tmp_path):
    name = tmp_path / "cp"
    restart.write(name, [np.array([1.0, np.nan, 3.0])], "binary")
    assert restart.read(name, "binary")[0].tolist() == [1.0, 0.0, 3.0]


def test_errors(tmp_path):
    with pytest.raises(FileNotFoundError):
        restart.read(tmp_path / "missing", "binary")
    (tmp_path / "short.bin").write_bytes(struct.pack("<ii", 1, 10) + b"\0" * 8)
    with pytest.raises(ValueError):
        restart.read(tmp_path / "short", "binary")
    with pytest.raises(ValueError):
        restart.write(tmp_path / "x", [[1.0]], "binary", real=np.float16)


def test_python_and_cpp_host_agree(tmp_path):
    data = [np.float32([1, 2, 3]), np.float32([4, 5, 6]), np.float32([0.123456789, -1e-7, 12345678.0, -0.0])]
    prefix = str(tmp_path / "rs")
    restart.write(prefix + "_py", data, "binary")
    restart.write(prefix + "_py", data, "ascii")
    restart.write(prefix + "_py64", [np.array([1.0, 2.0, 3.0]), np.array([0.1, -2.5e-300])], "binary",
                  real=np.float64)
    exe = build_host_test()
    r = subprocess.run([exe, "restart", prefix], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    for ext in (".bin", ".txt"):
        with open(prefix + "_cpp" + ext, "rb") as a, open(prefix + "_py" + ext, "rb") as b:
            assert a.read() == b.read(), ext
    with open(prefix + "_cpp64.bin", "rb") as a, open(prefix + "_py64.bin", "rb") as b:
        assert a.read() == b.read()
    assert not os.path.exists(prefix + "_del.bin")
