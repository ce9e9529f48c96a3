"""CPU check of the server's result-arena allocator (gpu-nurbs_amd/csrc/arena.hpp): a random
alloc/release sequence against a brute-force byte map -- no overlap, alignment, exact used
bytes, full coalescing back to one range, first-fit placement."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
#include "arena.hpp"
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>
int main() {
    const size_t cap = 1 << 20;
    gn::Arena a(cap);
    std::vector<unsigned char> owner(cap, 0);  // brute-force map, 256-byte granules
    struct R { size_t off, bytes; };
    std::vector<R> live;
    std::mt19937_64 rng(7);
    size_t used = 0;
    for (int step = 0; step < 200000; step++) {
        if (live.empty() || rng() % 3) {
            size_t b = rng() % 20000;
            size_t off;
            bool ok = a.alloc(b, &off);
            size_t need = gn::Arena::round(b ? b : 1);
            // brute force: does any free run of `need` bytes exist, and is off the first one?
            size_t first = cap, run = 0;
            for (size_t g = 0; g < cap; g += 256) {
                run = owner[g] ? 0 : run + 256;
                if (run == need) { first = g + 256 - need; break; }
            }
            if (ok != (first != cap)) { printf("FAIL fit %d\n", step); return 1; }
            if (!ok) continue;
            if (off != first || off % 256) { printf("FAIL first-fit %d\n", step); return 1; }
            for (size_t g = off; g < off + need; g += 256) {
                if (owner[g]) { printf("FAIL overlap %d\n", step); return 1; }
                owner[g] = 1;
            }
            used += need;
            live.push_back({off, b});
        } else {
            size_t k = rng() % live.size();
            R r = live[k];
            live[k] = live.back();
            live.pop_back();
            a.release(r.off, r.bytes);
            size_t need = gn::Arena::round(r.bytes ? r.bytes : 1);
            for (size_t g = r.off; g < r.off + need; g += 256) owner[g] = 0;
            used -= need;
        }
        if (a.used() != used) { printf("FAIL used %d\n", step); return 1; }
    }
    for (R r : live) a.release(r.off, r.bytes);
    if (a.used() != 0 || a.ranges() != 1 || a.largest_free() != cap) { printf("FAIL coalesce\n"); return 1; }
    if (a.peak() > cap) { printf("FAIL peak\n"); return 1; }
    printf("OK\n");
    return 0;
}
"""


def test_arena_against_brute_force(tmp_path):
    src = tmp_path / "arena_check.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "arena_check"
    inc = os.path.join(ROOT, "gpu-nurbs_amd", "csrc")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", inc, "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.strip() == "OK", out.stdout
