"""Scratch bounds of the network's deterministic gradient and loss reductions
(stroemung_amd/csrc/nn_sizes.h), compiled with g++ and checked on the CPU:
the buffer is sized once for the largest batch, so the partials of every
smaller batch must fit too, although the weight-gradient chunk count is not
monotone in the pixel count.  An overflow there would write past the
buffer on the GPU, so the bound is checked for every pixel count up to a
few million and on samples beyond, for every convolution of the network."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "stroemung_amd", "csrc")

DRIVER = r"""
#include <cstdio>
#include <cstdlib>
#include <random>
#include "nn_sizes.h"
using namespace stm::nn;

// (cin, cout) of the 14 convolutions of the U-Net
static const int CONV[][2] = {{32, 32}, {32, 32}, {32, 64}, {64, 64}, {64, 128}, {128, 128},
                              {128, 256}, {256, 256}, {384, 128}, {128, 128}, {192, 64},
                              {64, 64}, {96, 32}, {32, 32}};

static int fits(int64_t cap, int64_t npix) {
    for (auto& c : CONV) {
        const int64_t bound = grad_partials(c[0], c[1], cap);
        const int64_t ch = dw_chunks(npix);
        if (ch > 1 && ch * c[1] * 9 * c[0] > bound) {
            printf("FAIL dw cap=%lld npix=%lld chunks=%lld\n", (long long)cap, (long long)npix,
                   (long long)ch);
            return 0;
        }
        if (bias_blocks(npix) * c[1] > bound) {
            printf("FAIL bias cap=%lld npix=%lld\n", (long long)cap, (long long)npix);
            return 0;
        }
    }
    return 1;
}

int main(int argc, char** argv) {
    const int64_t dense = atoll(argv[1]);
    const int64_t caps[] = {64, 1023, 1024, 1025, 4096, 262145, 266239, 266240, 524288, 1 << 20,
                            3 << 20, 64LL << 20, (68LL << 20) + 5, 200LL << 20};
    std::mt19937_64 rng(7);
    for (int64_t cap : caps) {
        const int64_t top = cap < dense ? cap : dense;
        for (int64_t n = 1; n <= top; n++)
            if (!fits(cap, n)) return 1;
        for (int i = 0; i < 200000; i++)
            if (!fits(cap, 1 + (int64_t)(rng() % (uint64_t)cap))) return 1;
        if (!fits(cap, cap)) return 1;
    }
    // the loss and the head for nb <= cap samples of hw pixels
    const int64_t hws[] = {64, 4096, 16384, 65536};
    for (int64_t hw : hws)
        for (int cap = 1; cap <= 64; cap++) {
            const int64_t bound = loss_partials(hw, cap, 32);
            for (int nb = 1; nb <= cap; nb++)
                if (64 * nb > bound || blocks256(hw * nb) * 33 > bound) {
                    printf("FAIL loss hw=%lld cap=%d nb=%d\n", (long long)hw, cap, nb);
                    return 1;
                }
        }
    printf("ok\n");
    return 0;
}
"""


def _build(include_dir, out_dir):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    src, exe = out_dir / "sizes_check.cpp", out_dir / "sizes_check"
    src.write_text(DRIVER)
    subprocess.run([gxx, "-O2", "-std=c++17", "-I", str(include_dir), str(src), "-o", str(exe)],
                   check=True)
    return str(exe)


def test_partials_of_every_smaller_batch_fit_the_bound(tmp_path):
    exe = _build(CSRC, tmp_path)
    r = subprocess.run([exe, "3000000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr


def test_a_bound_of_256_chunks_is_caught(tmp_path):
    """The check is not vacuous: 256 instead of 257 chunks (the count reaches
    257 just above 256 chunk sizes) fails it."""
    text = open(os.path.join(CSRC, "nn_sizes.h")).read()
    bad = text.replace("std::max<int64_t>(257,", "std::max<int64_t>(256,")
    assert bad != text
    (tmp_path / "nn_sizes.h").write_text(bad)
    exe = _build(tmp_path, tmp_path)
    r = subprocess.run([exe, "300000"], capture_output=True, text=True, timeout=300)
    assert r.returncode != 0 and "FAIL dw" in r.stdout
