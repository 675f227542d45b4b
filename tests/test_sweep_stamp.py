"""The remesher's sweep stamps across the wrap of their 16-bit call tag
(samoa_amd/csrc/swe_sweep_stamp.hpp), compiled for the host with g++."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "samoa_amd", "csrc")

DRIVER = r"""
#include <cstdio>
#include <initializer_list>
#include "swe_sweep_stamp.hpp"
using namespace swe;
static int bad = 0;
#define CHECK(c) do { if (!(c)) { std::printf("FAIL line %d: %s\n", __LINE__, #c); ++bad; } } while (0)
int main() {
    // within one call: the previous sweep and this one are recent
    for (unsigned epoch : {1u, 7u, 0xfffeu, 0xffffu, 0x10000u, 0x10001u, 0x2ffffu}) {
        for (unsigned j = 2; j < 40; ++j) {
            const unsigned sw = sweep_number(epoch, j);
            CHECK(!stamp_is_old(sweep_number(epoch, j - 1), sw));
            CHECK(!stamp_is_old(sw, sw));
            if (j >= 3) CHECK(stamp_is_old(sweep_number(epoch, j - 2), sw));
            // every stamp of the previous few calls is old, whatever its index
            for (unsigned back = 1; back < 5; ++back)
                for (unsigned k : {1u, 2u, j - 1, j, j + 1, 0xfffeu})
                    CHECK(stamp_is_old(sweep_number(epoch - back, k), sw));
        }
    }
    // across the wrap: tag 0xffff, then tag 0
    CHECK(sweep_number(0xffffu, 3) == 0xffff0003u);
    CHECK(sweep_number(0x10000u, 3) == 0x00000003u);
    CHECK(stamp_is_old(sweep_number(0xffffu, 9), sweep_number(0x10000u, 2)));
    CHECK(stamp_is_old(sweep_number(0xfff0u, 0xfffe), sweep_number(0x10000u, 2)));
    // cleared stamps read as old from the second sweep of the wrapped call on
    CHECK(stamp_is_old(0u, sweep_number(0x10000u, 2)));
    CHECK(!stamp_is_old(0u, sweep_number(0x10000u, 1)));
    CHECK(stamp_tag_wraps(0x10000u) && stamp_tag_wraps(0x20000u) && stamp_tag_wraps(0u));
    CHECK(!stamp_tag_wraps(0x10001u) && !stamp_tag_wraps(0xffffu));
    if (!bad) std::printf("ok\n");
    return bad ? 1 : 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_stamps_across_tag_wrap(tmp_path):
    src = tmp_path / "stamp.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "stamp"
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{CSRC}", str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok"
