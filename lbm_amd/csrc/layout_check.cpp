// Host-only check of the buffer layouts (layout.h): for a set of grid shapes, storage sizes,
// lattices and boundary sets, in both the row-blocked and the plane arrangement,
//   * every (population, cell) of the padded grid has its own element, inside the buffer;
//   * the first interior cell of every row and population starts on a 128-byte boundary;
//   * per-cell indices are distinct and inside cplane;
//   * row-blocked buffers are chosen only while element offsets fit 31 bits and a slab of the
//     slowest axis is below 4 GiB, and every element of a slab lies within [slab start, slab
//     start + slab stride), the range of the step kernel's per-slab 32-bit offsets;
//   * the two launches of a split step (part_row) cover every cell of the halo axis exactly once
//     and map every other launch coordinate past the last cell, for any block height.
// Prints one JSON line per case; exits non-zero on the first violation.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "layout.h"

using namespace lbm;

static int check(int nx, int ny, int nz, int raw, int q, int pi, int pj, int pk, bool rowblock, int fast) {
    const Layout L = make_layout(nx, ny, nz, raw, pi, pj, pk, fast, q, rowblock);
    const long long cells = (long long)(nx + 2) * (ny + 2) * (nz + 2);
    if ((long long)q * cells > 400000000LL) return 0;   // keep the enumeration small
    std::vector<unsigned char> hit((size_t)L.total, 0), chit((size_t)L.cplane, 0);
    for (int k = 0; k <= nz + 1; ++k)
        for (int j = 0; j <= ny + 1; ++j)
            for (int i = 0; i <= nx + 1; ++i) {
                const long long c = L.cidx(i, j, k);
                if (c < 0 || c >= L.cplane || chit[(size_t)c]) return 2;
                chit[(size_t)c] = 1;
                for (int l = 0; l < q; ++l) {
                    const long long e = (long long)l * L.plane + L.cell(i, j, k);
                    if (e < 0 || e >= L.total || hit[(size_t)e]) return 1;
                    hit[(size_t)e] = 1;
                    int a0 = L.ax0 == 0 ? i : L.ax0 == 1 ? j : k;
                    if (a0 == 1 && (e * raw) % 128 != 0) return 3;
                    if (L.rowblock) {
                        const int a2 = L.ax2 == 0 ? i : L.ax2 == 1 ? j : k;
                        const long long s2 = L.ax2 == 0 ? L.si : L.ax2 == 1 ? L.sj : L.sk;
                        const long long w = e - (long long)a2 * s2;
                        if (w < 0 || w >= s2) return 9;
                    }
                }
            }
    if (L.rowblock && ((double)L.total >= 2147483648.0 ||
                       (double)(L.ax2 == 0 ? L.si : L.ax2 == 1 ? L.sj : L.sk) * raw >= 4294967296.0))
        return 4;
    std::printf("{\"dims\": [%d, %d, %d], \"raw\": %d, \"q\": %d, \"per\": [%d, %d, %d], \"rowblock\": %d, "
                "\"ax0\": %d, \"pitch\": %d, \"plane\": %lld, \"total\": %lld, \"cplane\": %lld}\n",
                nx, ny, nz, raw, q, pi, pj, pk, L.rowblock, L.ax0, L.pitch, L.plane, L.total, L.cplane);
    return 0;
}

int main() {
    const int dims[][3] = {{20, 14, 12}, {18, 13, 10}, {12, 8, 6}, {64, 24, 12}, {96, 32, 16}, {7, 5, 3}};
    const int pers[][3] = {{0, 1, 0}, {1, 0, 0}, {0, 0, 1}, {0, 0, 0}, {1, 1, 1}};
    int n = 0;
    for (const auto& d : dims)
        for (int raw : {2, 4, 8, 16})
            for (int q : {19, 27})
                for (const auto& p : pers)
                    for (int rb = 0; rb < 2; ++rb)
                        for (int fast = -1; fast < 3; ++fast) {
                            const int rc = check(d[0], d[1], d[2], raw, q, p[0], p[1], p[2], rb != 0, fast);
                            if (rc) {
                                std::printf("{\"error\": %d, \"dims\": [%d, %d, %d], \"raw\": %d, \"q\": %d, \"rowblock\": %d, \"fast\": %d}\n",
                                            rc, d[0], d[1], d[2], raw, q, rb, fast);
                                return rc;
                            }
                            ++n;
                        }
    // the size rule: D3Q27 FP32 at 768x256x192 (5.2 GB per buffer) stays
    // row-blocked; a grid with more than 2^31 elements per buffer does not
    const Layout big = make_layout(768, 256, 192, 4, 0, 1, 0, -1, 27, true);
    const Layout mid = make_layout(384, 192, 128, 4, 0, 1, 0, -1, 27, true);
    const Layout huge = make_layout(2048, 1024, 1024, 2, 0, 1, 0, -1, 27, true);
    if (!big.rowblock || !mid.rowblock || huge.rowblock) return 5;
    // split steps: launch heights rounded up to the block height, as the launcher's grid is
    int nsplit = 0;
    for (int e = 1; e <= 2; ++e)
    for (int len = 2 * e + 1; len <= 1024; ++len)
        for (int by : {1, 2, 4, 8}) {
            std::vector<int> cnt(len, 0);
            for (int part = 1; part <= 2; ++part) {
                const int rows = part == 1 ? 2 * e : len - 2 * e;
                const int launched = (rows + by - 1) / by * by;
                for (int c = 0; c < launched; ++c) {
                    const int m = part_row(c, part, len, e);
                    if (m == len) {
                        if (c < rows) return 6;   // a coordinate of the part mapped outside
                        continue;
                    }
                    if (m < 0 || m > len || c >= rows) return 7;
                    ++cnt[m];
                }
            }
            for (int x = 0; x < len; ++x)
                if (cnt[x] != 1) return 8;
            ++nsplit;
        }
    std::printf("{\"cases\": %d, \"rowblock_768x256x192_f32\": %d, \"rowblock_384x192x128_f32\": %d, "
                "\"rowblock_2048x1024x1024_f16\": %d, \"split_heights\": %d}\n", n, big.rowblock, mid.rowblock,
                huge.rowblock, nsplit);
    return 0;
}
