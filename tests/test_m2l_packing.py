"""Block-diagonal packing of the M2L operator blocks into 16x4 MFMA tiles
(simplefmm_amd/csrc/m2l_sched.hpp), checked on the host for every order 1..16.

A small C++ program built from the header enumerates, for each packed group,
the (block, row, column) operator entries its tiles hold and checks that every
entry of every block of stages 1, 2 and 3 is covered exactly once, that no
group exceeds 16 rows or 8 pieces, that k-steps cover the group's columns,
that the fragment counts fit the strides the kernel and host use, and that the
wave schedule assigns every group exactly once.  The GPU parity tests check the
packed kernels' results against the reference and the oracle.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))

PROGRAM = r"""
#include "m2l_sched.hpp"
#include <cstdio>
#include <map>
#include <tuple>
using namespace sfmm;

static int fail(const char* what, int P, int st) { std::printf("FAIL %s P=%d stage=%d\n", what, P, st); return 1; }

// stage 1: cos block n (n+1 x n+1), sin block n (n x n); stage 2: block k (P+1-k square) once per part
static int check12(int P, int st)
{
    const Packing pk = pack_stage(P, st);
    std::map<std::tuple<int, int, int, int>, int> seen;  // (n, part, row, col) -> count
    for (int g = 0; g < pk.ng; ++g) {
        const Group& G = pk.g[g];
        if (G.rows > 16 || G.np > kMaxPieces) return fail("group too large", P, st);
        if (G.nks() * 4 < G.cols) return fail("k-steps short of the columns", P, st);
        for (int i = 0; i < 16; ++i)
            for (int f = 0; f < 4 * G.nks(); ++f) {
                const int qr = pk.piece_of_row(g, i), qc = pk.piece_of_col(g, f);
                if (qr < 0 || qr != qc) continue;
                const Piece& pc = G.pc[qr];
                seen[{pc.n, pc.part, pc.r0 + i - G.row0[qr], f - G.col0[qr]}]++;
            }
    }
    for (int n = 0; n <= P; ++n)
        for (int part = 0; part < 2; ++part) {
            const int size = st == 1 ? (part == 0 ? n + 1 : n) : ((part == 1 && n == 0) ? 0 : P + 1 - n);
            for (int r = 0; r < size; ++r)
                for (int c = 0; c < size; ++c) {
                    auto it = seen.find({n, part, r, c});
                    if (it == seen.end() || it->second != 1) return fail("entry not covered exactly once", P, st);
                    seen.erase(it);
                }
        }
    if (!seen.empty()) return fail("entry outside every block", P, st);
    const int stride = st == 1 ? rot_nfrag(P) : pk.nfrag();
    if (pk.nfrag() > stride) return fail("fragments exceed the stride", P, st);
    return 0;
}

// stage 3: degree j, cos tile rows a = 0..j cols b = 0..j, sin tile rows a = 1..j cols b = 1..j,
// both tiles of a group with the same row layout
static int check3(int P)
{
    const Packing pk = pack_stage3(P);
    std::map<std::tuple<int, int, int, int>, int> seen;
    for (int g = 0; g < pk.ng; ++g) {
        const Group& G = pk.g[g];
        if (G.rows > 16 || G.np > kMaxPieces) return fail("group too large", P, 3);
        if (G.nks() * 4 < G.cols || G.nks_s() * 4 < G.cols_s) return fail("k-steps short", P, 3);
        for (int part = 0; part < 2; ++part)
            for (int i = 0; i < 16; ++i)
                for (int f = 0; f < 4 * (part ? G.nks_s() : G.nks()); ++f) {
                    const int qr = pk.piece_of_row(g, i);
                    const int qc = part ? pk.piece_of_col_s(g, f) : pk.piece_of_col(g, f);
                    if (qr < 0 || qr != qc) continue;
                    const Piece& pc = G.pc[qr];
                    const int a = pc.r0 + i - G.row0[qr];
                    const int b = part ? f - G.col0s[qr] + 1 : f - G.col0[qr];
                    if (part == 1 && a == 0) continue;  // the sin row of order 0 is zero
                    seen[{pc.n, part, a, b}]++;
                }
    }
    for (int j = 0; j <= P; ++j)
        for (int part = 0; part < 2; ++part)
            for (int a = part; a <= j; ++a)
                for (int b = part; b <= j; ++b) {
                    auto it = seen.find({j, part, a, b});
                    if (it == seen.end() || it->second != 1) return fail("entry not covered exactly once", P, 3);
                    seen.erase(it);
                }
    if (!seen.empty()) return fail("entry outside every block", P, 3);
    if (pk.frag3(pk.ng) > rot_nfrag(P)) return fail("fragments exceed the stride", P, 3);
    return 0;
}

template <int P, int NW, bool PIPE>
static int check_schedule()
{
    constexpr M2LSchedule<P, NW, PIPE> S{};
    const int ng[3] = {pack_stage(P, 1).ng, pack_stage(P, 2).ng, pack_stage3(P).ng};
    const WaveTiles* T[3] = {&S.s1, &S.s2, &S.s3};
    for (int s = 0; s < 3; ++s) {
        int count[64] = {};
        for (int w = 0; w < NW; ++w)
            for (int q = 0; q < T[s]->count[w]; ++q) count[T[s]->item[w][q]]++;
        for (int g = 0; g < ng[s]; ++g)
            if (count[g] != 1) return fail("schedule does not assign a group exactly once", P, s + 1);
    }
    return 0;
}

int main()
{
    int bad = 0;
    for (int P = 1; P <= 16; ++P) bad |= check12(P, 1) | check12(P, 2) | check3(P);
    bad |= check_schedule<4, 12, true>() | check_schedule<8, 12, true>() | check_schedule<12, 12, true>();
    bad |= check_schedule<13, 8, false>() | check_schedule<16, 8, false>();
    const Packing p1 = pack_stage(12, 1), p2 = pack_stage(12, 2), p3 = pack_stage3(12);
    std::printf("p12 fragments %d %d %d\n", p1.nfrag(), p2.nfrag(), p3.frag3(p3.ng));
    return bad;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_packing_covers_every_operator_entry_once(tmp_path):
    src = tmp_path / "pack.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "pack"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "simplefmm_amd", "csrc"), str(src), "-o",
                    str(exe)], check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    # p = 12: 43 + 43 + 45 MFMAs per entry instead of 52 + 52 + 52
    assert "p12 fragments 43 43 45" in r.stdout


@pytest.mark.parametrize("p", [1, 4, 8, 12])
def test_dense_m2l_operators_match_staged_rotations(tmp_path, p):
    # the f16 hi/lo class operators of the mixed M2L (with their azimuth groups and
    # the dense index order) reproduce rotate-translate-rotate applied stage by
    # stage, for every interaction offset, to the 22 bits the split carries
    exe = tmp_path / "dense_ops_check"
    lib = os.path.join(ROOT, "simplefmm_amd")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(lib, "csrc"),
                    os.path.join(ROOT, "tests", "dense_ops_check.cpp"), "-o", str(exe), "-L", lib, "-lsfmm",
                    "-Wl,-rpath," + lib], check=True, timeout=300)
    out = subprocess.run([str(exe), str(p)], capture_output=True, text=True, timeout=300, check=True).stdout
    tok = out.split()
    assert tok[0] == "max_rel" and int(tok[3]) == 316, out
    assert float(tok[1]) < 2e-6, out
