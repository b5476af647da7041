"""Key-table size rules (joinproject-dim3_amd/csrc/dim3_sizes.h), compiled on
the CPU with g++: every table the API can build for an accepted relation
(at most 2^30 - 1 tuples) has at most 2^31 slots, so its slot count plus the
slot of the key -1 fits 32 bits; value ranges whose power-of-two slot space
would pass 2^31 are hashed instead; the estimate's table is bounded by the
int32 values its sample can hold."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROG = r'''
#include <stdio.h>
#include "dim3_sizes.h"
using namespace dim3s;
#define CHECK(c) do { if (!(c)) { printf("FAIL %s\n", #c); return 1; } } while (0)
int main() {
    CHECK(pow2_at_least(0) == 64 && pow2_at_least(64) == 64 && pow2_at_least(65) == 128);
    CHECK(pow2_at_least((1ull << 31) + 1) == (1ull << 32));  // 64-bit: no truncation to 0
    CHECK(key_table_cap(kMaxTuples) == kMaxSlots);
    CHECK(key_table_cap(kMaxTuples) + 1 <= 0xFFFFFFFFull);
    CHECK(key_table_cap(1ull << 29) == (1ull << 30));
    CHECK(key_table_cap((1ull << 30) + 1) == (1ull << 32));  // past 2^30 tuples: 2^32 slots
    CHECK(range_fits(100, 50) && !range_fits(165, 50));
    CHECK(range_fits(1ull << 31, kMaxTuples));
    CHECK(!range_fits((1ull << 31) + 1, kMaxTuples));       // would need 2^32 slots: hashed
    CHECK(estimate_table_cap(5000, 0) == 16384);
    CHECK(estimate_table_cap(kMaxTuples, 17) == key_table_cap(1ull << 15));
    CHECK(estimate_table_cap(kMaxTuples, 32) == 64);
    printf("ok\n");
    return 0;
}
'''


def test_size_rules(tmp_path):
    src = tmp_path / "sizes.cpp"
    src.write_text(PROG)
    exe = str(tmp_path / "sizes")
    subprocess.run(["g++", "-std=c++14", "-O1", "-I" + os.path.join(ROOT, "joinproject-dim3_amd", "csrc"), "-o", exe,
                    str(src)], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout
