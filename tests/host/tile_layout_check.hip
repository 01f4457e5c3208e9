// Host check of the fused stage kernel's tile bookkeeping (stage_kernel.hpp), for every
// tile shape the kernels are built for: the compile-time face-run table covers every x
// and y face of the tile exactly once, every LDS read of phase 2 stays inside the
// primitive array (and out of the pad column of its odd row pitch), every conserved-flux slot of the trimmed layout is written exactly
// once, and the reads of phase 3 (flux difference, corner EMFs, including the open
// lo-side cases) stay inside what phase 2 wrote.  Built and run by
// tests/test_tile_layout.py; exit status 0 when every check holds.
#include <cstdio>
#include <vector>
#include "stage_kernel.hpp"
using namespace chimera;

template <int TX, int TY>
int check() {
    using S = TileShape<TX, TY>;
    constexpr FaceRunTable<TX, TY> T{};
    int err = 0;
    std::vector<int> fxs(S::FX_SZ, 0), fys(S::FY_SZ, 0), c4x(S::FX4_SZ, 0), c4y(S::FY4_SZ, 0);
    for (int t = 0; t < TX * TY; ++t) {
        const FaceRun run = T.r[t];
        const int dir = run.dir;
        const int sA = dir ? S::PW : 1, sT = dir ? 1 : S::PW;
        // inside the array and never in the pad column of the odd row pitch
        auto inP = [&](int p) { if (p < 0 || p >= S::P_SZ || p % S::PW >= S::LW) { ++err; } };
        if (run.len > 0) {
            const int fx0 = dir ? run.line : run.start, fy0 = dir ? run.start : run.line;
            const int lpos0 = (fy0 + 1) * S::PW + (fx0 + 1);
            inP(lpos0 - sA); inP(lpos0); inP(lpos0 + sA);
        }
        for (int k = 0; k < run.len; ++k) {
            const int fx = dir ? run.line : run.start + k;
            const int fy = dir ? run.start + k : run.line;
            const int lpos = (fy + 1) * S::PW + (fx + 1);
            inP(lpos - sA); inP(lpos); inP(lpos + sA); inP(lpos + 2 * sA); inP(lpos + sT); inP(lpos - sT);
            if (dir) {
                if (fx < 0 || fx >= S::FYW || fy < 0 || fy >= S::FYH) ++err; else fys[fy * S::FYW + fx]++;
                if (fx >= 1 && fx <= TX) { int i4 = fy * TX + (fx - 1); if (i4 < 0 || i4 >= S::FY4_SZ) ++err; else c4y[i4]++; }
            } else {
                if (fx < 0 || fx >= S::FXW || fy < 0 || fy >= S::FXH) ++err; else fxs[fy * S::FXW + fx]++;
                if (fy >= 1 && fy <= TY) { int i4 = (fy - 1) * S::FXW + fx; if (i4 < 0 || i4 >= S::FX4_SZ) ++err; else c4x[i4]++; }
            }
        }
    }
    for (int v : fxs) if (v != 1) ++err;
    for (int v : fys) if (v != 1) ++err;
    for (int v : c4x) if (v != 1) ++err;
    for (int v : c4y) if (v != 1) ++err;
    // phase 3 reads
    for (int cj = 0; cj < TY; ++cj)
        for (int ci = 0; ci < TX; ++ci) {
            int xr = cj * S::FXW + (ci + 1), yr = (cj + 1) * TX + ci;
            if (xr < 0 || xr >= S::FX4_SZ || xr - 1 < 0) ++err;
            if (yr >= S::FY4_SZ || yr - TX < 0) ++err;
            // the flux difference reads exactly the faces of the cell
            if (xr - 1 != cj * S::FXW + ci) ++err;
            int ab[5][2] = {{ci, cj}, {ci, cj - 1}, {ci - 1, cj}, {ci, cj + 1}, {ci + 1, cj}};
            for (int n = 0; n < 5; ++n) {
                if (n == 3 && cj != 0) continue;   // negEz(ci, cj+1): only with an open y-lo side
                if (n == 4 && ci != 0) continue;   // negEz(ci+1, cj): only with an open x-lo side
                int a = ab[n][0], b = ab[n][1];
                int i1 = (b + 1) * S::FXW + (a + 1), i2 = (b + 2) * S::FXW + (a + 1);
                int i3 = (b + 1) * S::FYW + (a + 1), i4 = (b + 1) * S::FYW + (a + 2);
                if (i1 < 0 || i1 >= S::FX_SZ || i2 < 0 || i2 >= S::FX_SZ) ++err;
                if (i3 < 0 || i3 >= S::FY_SZ || i4 < 0 || i4 >= S::FY_SZ) ++err;
            }
        }
    std::printf("%dx%d: LDS %d bytes, errors %d\n", TX, TY, S::LDS_DOUBLES * 8, err);
    return err;
}

int main() {
    int e = check<32, 8>() + check<16, 16>() + check<64, 4>() + check<32, 32>() + check<32, 16>();
    return e ? 1 : 0;
}
