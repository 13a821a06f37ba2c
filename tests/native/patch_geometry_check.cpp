// Host check of the one-launch first sweep's patch geometry (PatchCol,
// patch_red / patch_black / patch_black_nbrs in orbit_internal.h), over every
// patch of many grids and five patch shapes: every grid column is owned by
// exactly one patch, red items are red and black items black, no two red items
// of a patch share an ext position, every black column's existing neighbours
// (west_col / east_col, j +- 1, as the sweeps define them) are the red items at
// the positions patch_black_nbrs names, and every index is in range.  Built
// and run by tests/test_patch_geometry.py; prints the number of problems.
#include "orbit_internal.h"
#include <cstdio>
#include <vector>
using namespace orbit;
template <int PA, int PB>
int check(int nx, int ny, int px) {
    const int npx = (nx + PA - 1) / PA, npy = (ny + PB - 1) / PB;
    constexpr int NR = PA * PB / 2 + PA + PB, NB = PA * PB / 2, EW = PA + 2, NPOS = (PA + 2) * (PB + 2);
    std::vector<int> owned(nx * ny, 0);
    int bad = 0;
    for (int pj = 0; pj < npy; ++pj)
        for (int pi = 0; pi < npx; ++pi) {
            const int i0 = pi * PA, j0 = pj * PB;
            std::vector<int> poscol(NPOS, -2);
            for (int t = 0; t < NR; ++t) {
                PatchCol r = patch_red<PA, PB>(t, i0, j0, nx, ny, px);
                if (r.col < 0 || r.col >= nx * ny || r.pos < 0 || r.pos >= NPOS) { printf("red idx out of range\n"); ++bad; }
                if (!r.valid) continue;
                if (((r.i + r.j) & 1) != 0) { printf("red item not red nx%d ny%d t%d\n", nx, ny, t); ++bad; }
                if (poscol[r.pos] != -2) { printf("pos twice\n"); ++bad; }
                poscol[r.pos] = r.col;
                if (t < PA * PB / 2) owned[r.col]++;
            }
            for (int t = 0; t < NB; ++t) {
                PatchCol r = patch_black<PA, PB>(t, i0, j0, nx, ny);
                if (r.col < 0 || r.col >= nx * ny || r.pos < 0 || r.pos >= NPOS) { printf("black idx out of range\n"); ++bad; }
                if (!r.valid) continue;
                if (((r.i + r.j) & 1) != 1) { printf("black item not black\n"); ++bad; }
                owned[r.col]++;
                int q[4];
                patch_black_nbrs<PA>(r.pos, r.i, i0, nx, q);
                const int nb[4] = {west_col(r.col, r.i, nx, px), east_col(r.col, r.i, nx, px),
                                   r.j > 0 ? r.col - nx : -1, r.j < ny - 1 ? r.col + nx : -1};
                for (int d = 0; d < 4; ++d) {
                    if (q[d] < 0 || q[d] >= NPOS) { printf("nbr pos out of range\n"); ++bad; continue; }
                    if (nb[d] < 0) continue;
                    if (poscol[q[d]] != nb[d]) {
                        printf("PA%d PB%d nx%d ny%d px%d patch(%d,%d) black (%d,%d) dir %d: pos %d holds %d want %d\n", PA, PB,
                               nx, ny, px, pi, pj, r.i, r.j, d, q[d], poscol[q[d]], nb[d]);
                        ++bad;
                    }
                }
            }
        }
    for (int c = 0; c < nx * ny; ++c)
        if (owned[c] != 1) { printf("PA%d PB%d nx%d ny%d px%d: column %d owned %d times\n", PA, PB, nx, ny, px, c, owned[c]); ++bad; }
    return bad;
}
int main() {
    int bad = 0, n = 0;
    const int nxs[] = {2, 4, 6, 8, 10, 12, 24, 48, 64, 66, 256, 3, 5, 7, 9, 11, 25};
    const int nys[] = {1, 2, 3, 4, 5, 6, 7, 8, 20, 40, 70, 71, 200};
    for (int nx : nxs)
        for (int ny : nys)
            for (int px = 0; px < 2; ++px) {
                if (px && (nx & 1)) continue;  // two colours only
                bad += check<4, 4>(nx, ny, px) + check<8, 4>(nx, ny, px) + check<4, 2>(nx, ny, px) +
                       check<2, 2>(nx, ny, px) + check<8, 8>(nx, ny, px);
                n += 5;
            }
    printf("%d geometries checked, %d problems\n", n, bad);
    return bad != 0;
}
