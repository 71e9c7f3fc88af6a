"""Host build of the shared topology header (hp3d_amd/csrc/hp3d_tables.h):
the branch-free descriptor scan the fused kernels use (desc_h1_scan) gives
the same H1 dof descriptors as the hp3D-ordered walk (desc_h1, pinned by the
GPU parity tests against the oracle), over random anisotropic orders and all
edge and face orientation codes.  Runs on the CPU with g++."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "..", "hp3d_amd", "csrc", "hp3d_tables.h")

SRC = r"""
#define __host__
#define __device__
#include <cstdio>
#include <cstdlib>
#include "%s"
using namespace hp3d;
int main(int argc, char **argv)
{
    const int trials = atoi(argv[1]);
    srand(7);
    long bad = 0, tot = 0;
    for (int it = 0; it < trials; ++it) {
        int hdr[37];
        for (int e = 0; e < 12; ++e) hdr[e] = 1 + rand() %% MAXP;
        for (int f = 0; f < 6; ++f) hdr[12 + f] = (1 + rand() %% MAXP) * MODORDER + 1 + rand() %% MAXP;
        hdr[18] = (1 + rand() %% MAXP) * 100 + (1 + rand() %% MAXP) * 10 + 1 + rand() %% MAXP;
        for (int e = 0; e < 12; ++e) hdr[19 + e] = rand() %% 2;
        for (int f = 0; f < 6; ++f) hdr[31 + f] = rand() %% 8;
        ElemOrders o;
        if (!decode_orders(hdr, hdr + 19, hdr + 31, o)) return 2;
        int n, nb;
        dof_counts(0, o, n, nb);
        for (int k = 0; k < n; ++k) {
            ++tot;
            bad += desc_h1(k, o) != desc_h1_scan(k, o);
        }
    }
    printf("%%ld %%ld\n", tot, bad);
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_descriptor_scan_matches_walk(tmp_path):
    src = tmp_path / "scan.cpp"
    src.write_text(SRC % os.path.abspath(HEADER))
    exe = tmp_path / "scan"
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe), "20000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    tot, bad = map(int, r.stdout.split())
    assert tot > 1_000_000 and bad == 0, (tot, bad)
