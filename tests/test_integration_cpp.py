"""The src/cpu binding of INTEGRATION.md section 1, compiled as a reference
maintainer would: the snippet is taken from the document itself, wrapped in a
main() that declares the reference's Element struct
(cpu_mkl_baseline.cpp:16-19) and driver variables (:310-330), compiled with
g++ against include/acceles.h (its static_assert included) and linked against
libacceles.so.  Without a GPU the binary must read the matrix and then stop at
the upload with the library's message; on a GPU it runs top_k_spmv and prints
the reference's three timing lines."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PRELUDE = """#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>
struct Element {  // src/cpu/cpu_mkl_baseline.cpp:16-19
    float value;
    int index;
};
"""


def snippet() -> str:
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    sec = text[text.index("## 1."):text.index("## 2.")]
    m = re.search(r"```cpp\n(.*?)```", sec, re.S)
    assert m, "INTEGRATION.md section 1 has no C++ snippet"
    return m.group(1)


def driver_source() -> str:
    code = snippet()
    lines = code.splitlines()
    split = next(i for i, l in enumerate(lines) if l.startswith("for ("))
    file_scope, body = "\n".join(lines[:split]), "\n".join(lines[split:])
    return (PRELUDE + file_scope + "\n\nint main(int argc, char **argv) {\n"
            "    std::vector<std::string> input_paths(argv + 1, argv + argc);\n"
            "    int k = 100, num_threads = 1, num_test = 3;\n" + body + "\n    return 0;\n}\n")


def build(tmp_path) -> str:
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    lib = os.path.join(ROOT, "acceles_amd")
    assert os.path.exists(os.path.join(lib, "libacceles.so")), "build the library first"
    src = tmp_path / "cpu_driver.cpp"
    src.write_text(driver_source())
    exe = str(tmp_path / "cpu_driver")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-o", exe, str(src), "-I", os.path.join(ROOT, "include"),
                    "-L", lib, "-lacceles", f"-Wl,-rpath,{lib}"], check=True, capture_output=True, text=True)
    return exe


def write_matrix(acc, path, rows=3000, cols=64):
    csr = acc.CSRMatrix.generate(rows, cols, deg=(5, 20), seed=77)
    ptr, idx, val = csr.arrays()
    row = np.repeat(np.arange(rows, dtype=np.int32), np.diff(ptr))
    acc.COOMatrix.from_arrays(rows, cols, row, idx, val).write_mtx(str(path))
    return csr


def test_src_cpu_binding_compiles_and_links(acc, tmp_path):
    exe = build(tmp_path)
    mtx = tmp_path / "m.mtx"
    write_matrix(acc, mtx)
    r = subprocess.run([exe, str(mtx)], capture_output=True, text=True, timeout=120)
    if acc.device_count() == 0:
        assert r.returncode == 1
        assert "acc_gpu_upload" in r.stderr and "no GPU visible" in r.stderr, r.stderr
    else:
        assert r.returncode == 0, r.stderr
        for name in ("compute time:", "sort time:", "exec time:"):
            assert name in r.stdout


def test_src_cpu_binding_reports_a_missing_file(acc, tmp_path):
    exe = build(tmp_path)
    r = subprocess.run([exe, str(tmp_path / "absent.mtx")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "acc_read_coo_matrix" in r.stderr


@pytest.mark.gpu
def test_src_cpu_binding_runs_on_the_gpu(acc, gpu, tmp_path):
    exe = build(tmp_path)
    mtx = tmp_path / "m.mtx"
    write_matrix(acc, mtx, rows=50000, cols=500)
    r = subprocess.run([exe, str(mtx), str(mtx)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("exec time:") == 2
