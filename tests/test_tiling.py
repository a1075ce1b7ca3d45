"""The blocked red-black solver's tiling of a channel axis
(stroemung_amd/csrc/tiling.h), compiled with g++ and checked on the CPU for
every compiled variant's region and halo: the parts cover the axis in order;
a part's output has H cells of halo inside its region on every side that is
not the channel's edge, and on an edge side the channel's first / last cell
is an inner cell of the region (an exact tile needs nothing beyond the
edge); starts and region origins keep y pairs aligned.  The GPU tests check
the same tiling for bit-exact sweeps against the oracle."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "stroemung_amd", "csrc")

DRIVER = r"""
#include <cstdio>
#include <cstdlib>
#include "tiling.h"
using stm::plan_axis;

// E region cells, H halo, q alignment: every n in [3, nmax]
static int check(int E, int H, int q, int nmax) {
    for (int n = 3; n <= nmax; n++) {
        const auto v = plan_axis(n, E, H, q);
        int pos = 0;
        for (size_t i = 0; i < v.size(); i++) {
            const auto p = v[i];
            const int end = p.start + p.len, r0 = p.origin, r1 = p.origin + E;
            const bool cover = p.start == pos && p.len >= 1;
            const bool align = p.start % q == 0 && ((p.origin % q) + q) % q == 0;
            const bool left = (p.start - r0 >= H) || (p.start == 0 && -r0 >= 1);
            const bool right = (end + H <= r1) || (end == n && n - 1 <= r1 - 2);
            if (!cover || !align || !left || !right) {
                printf("FAIL E=%d H=%d q=%d n=%d part %zu (%d, %d, %d)\n", E, H, q, n, i,
                       p.start, p.len, p.origin);
                return 1;
            }
            pos = end;
        }
        if (pos != n) {
            printf("FAIL E=%d H=%d q=%d n=%d: covers %d\n", E, H, q, n, pos);
            return 1;
        }
    }
    return 0;
}

int main(int argc, char** argv) {
    // argv: nmax, then K EX EY per variant
    const int nmax = atoi(argv[1]);
    for (int a = 2; a + 2 < argc; a += 3) {
        const int K = atoi(argv[a]), EX = atoi(argv[a + 1]), EY = atoi(argv[a + 2]);
        const int H = 2 * K + 4;
        if (check(EX, H, 1, nmax) || check(EY, H, 2, nmax)) return 1;
        printf("K=%d region %dx%d: x %zu parts, y %zu parts for 1026 cells\n", K, EX, EY,
               plan_axis(1026, EX, H, 1).size(), plan_axis(1026, EY, H, 2).size());
    }
    return 0;
}
"""


def variants():
    """(K, EX, EY) of STM_TB_VARIANTS in kernels_rbtb.hip: V(i, K, R, HY, NS)."""
    text = open(os.path.join(CSRC, "kernels_rbtb.hip")).read()
    block = text[text.index("#define STM_TB_VARIANTS(V)"):]
    block = block[:block.index("\n\n")]
    out = []
    for i, k, r, hy, ns in re.findall(r"V\((\d+), (\d+), (\d+), (\d+), (\d+)\)", block):
        out.append((int(k), int(ns) * int(r), 2 * int(hy)))
    return out


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    d = tmp_path_factory.mktemp("tiling")
    src, exe = d / "tiling_check.cpp", d / "tiling_check"
    src.write_text(DRIVER)
    subprocess.run([gxx, "-O2", "-std=c++17", "-I", CSRC, str(src), "-o", str(exe)],
                   check=True)
    return str(exe)


def test_variant_list_is_parsed():
    v = variants()
    assert len(v) >= 10 and (6, 96, 96) in v and (8, 96, 128) in v


def test_axis_tiling_invariants_for_every_variant(driver):
    args = [str(x) for kxy in variants() for x in kxy]
    r = subprocess.run([driver, "2100"] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    # the bench grid (1024 interior + ring = 1026 cells a side): 16 x 16 tiles of
    # the K = 6 variant, one per CU
    assert "K=6 region 96x96: x 16 parts, y 16 parts for 1026 cells" in r.stdout


def test_a_missing_edge_allowance_is_caught(driver, tmp_path):
    """The check is not vacuous: a planner whose first part ignores the
    region's one-cell start outside the channel (E - H instead of E - 1 - H
    cells in x) fails it."""
    bad = open(os.path.join(CSRC, "tiling.h")).read().replace(
        "e0 = E - q - H", "e0 = E - H")
    assert bad != open(os.path.join(CSRC, "tiling.h")).read()
    (tmp_path / "tiling.h").write_text(bad)
    src, exe = tmp_path / "t.cpp", tmp_path / "t"
    src.write_text(DRIVER)
    subprocess.run([shutil.which("g++"), "-O2", "-std=c++17", "-I", str(tmp_path), str(src),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), "400", "6", "96", "96"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode != 0 and "FAIL" in r.stdout
