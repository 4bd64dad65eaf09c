"""Memory admission of the serving loop (acc::Budget, used by
acc_server_add_matrix): concurrent callers reserve, resize and release device
bytes against one limit.  Compiled with g++ against the library's internal
header and run on the CPU with many threads: the limit is never exceeded, and
exactly as many matrices are admitted as fit."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SOURCE = r"""
#include "internal.hpp"
#include <atomic>
#include <cstdio>
#include <thread>
#include <vector>

int main() {
    // a matrix is reserved at an estimate, then resized to its measured bytes
    // (as acc_server_add_matrix does around the upload and warm-up)
    const int64_t limit = 1000, est = 90, actual = 110;
    acc::Budget b;
    b.set_limit(limit);
    std::atomic<int> admitted{0}, refused{0};
    std::atomic<int64_t> peak{0};
    std::vector<std::thread> th;
    for (int t = 0; t < 32; ++t)
        th.emplace_back([&] {
            for (int rep = 0; rep < 200; ++rep) {
                if (!b.reserve(est)) { ++refused; continue; }
                int64_t u = b.used();
                int64_t p = peak.load();
                while (u > p && !peak.compare_exchange_weak(p, u)) {}
                if (!b.resize(est, actual)) { b.release(est); ++refused; continue; }
                u = b.used();
                p = peak.load();
                while (u > p && !peak.compare_exchange_weak(p, u)) {}
                ++admitted;
                if (rep % 2) b.release(actual);  // some matrices stay resident, others are removed
            }
        });
    for (auto &x : th) x.join();
    const int64_t resident = b.used();
    std::printf("admitted %d refused %d peak %lld resident %lld\n", admitted.load(), refused.load(),
                (long long)peak.load(), (long long)resident);
    if (peak.load() > limit) return 1;              // never above the limit
    if (resident % actual != 0) return 2;           // no reservation leaked
    if (resident / actual != limit / actual) return 3;  // as many resident matrices as fit
    if (b.reserve(limit - resident + 1)) return 4;  // one byte more than the room is refused
    if (!b.reserve(limit - resident)) return 5;     // exactly the room is admitted
    return 0;
}
"""


def test_budget_admission_concurrent(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = tmp_path / "budget.cpp"
    src.write_text(SOURCE)
    exe = str(tmp_path / "budget")
    subprocess.run(["g++", "-std=c++17", "-O1", "-pthread", "-I", os.path.join(ROOT, "acceles_amd", "csrc"),
                    "-o", exe, str(src)], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
