// tilemap_check -- host check of the strided passes' tile order
// (tdda_tilemap.hpp): every map is a bijection onto (row, block) -- for the
// x pass (row, block, right-hand side) of a group of R -- and in every full
// group of 8 rows (8 mirrored row pairs for the x pass) all tiles of one row
// (pair, in every right-hand side) lie on one residue mod 8, i.e. on one XCD.  Exit 0 when
// every case holds; prints the first failure otherwise.
#include <cstdio>
#include <vector>

#include "tdda_tilemap.hpp"

using namespace tdda;

int main()
{
    long cases = 0;
    for (int nrows = 1; nrows <= 70; nrows++)
        for (int nfb = 1; nfb <= 13; nfb++) {
            std::vector<int> seen((size_t)nrows * nfb, 0), res(nrows, -1);
            for (int r = 0; r < nrows * nfb; r++) {
                int o, fb;
                ytile_map(r, nrows, nfb, o, fb);
                if (o < 0 || o >= nrows || fb < 0 || fb >= nfb || seen[(size_t)o * nfb + fb]++) {
                    std::printf("y map not a bijection: nrows %d nfb %d r %d -> (%d, %d)\n", nrows, nfb, r, o, fb);
                    return 1;
                }
                if ((o / 8 + 1) * 8 <= nrows) {
                    if (res[o] < 0) res[o] = r % 8;
                    else if (res[o] != r % 8) {
                        std::printf("y row %d split over XCDs: nrows %d nfb %d\n", o, nrows, nfb);
                        return 1;
                    }
                }
            }
            cases++;
        }
    for (int R : {1, 2, 3, 8})
        for (int My = 2; My <= 300; My += (R == 1 ? 1 : 7))
            for (int nfb = 1; nfb <= 13; nfb++) {
                const int npair = (My - 1) / 2;
                const int T = My * nfb;  // tiles of one right-hand side
                std::vector<int> seen((size_t)T * R, 0), res(My, -1);
                for (int r = 0; r < T * R; r++) {
                    int ky, fb, k;
                    xtile_map(r, My, nfb, R, ky, fb, k);
                    if (ky < 0 || ky >= My || fb < 0 || fb >= nfb || k < 0 || k >= R ||
                        seen[((size_t)k * My + ky) * nfb + fb]++) {
                        std::printf("x map not a bijection: My %d nfb %d R %d r %d -> (%d, %d, %d)\n", My, nfb, R, r,
                                    ky, fb, k);
                        return 1;
                    }
                    const int pair = (ky >= 1 && ky <= npair) ? ky - 1 : (ky >= My - npair ? My - 1 - ky : -1);
                    if (pair >= 0 && (pair / 8 + 1) * 8 <= npair) {
                        if (res[pair] < 0) res[pair] = r % 8;
                        else if (res[pair] != r % 8) {
                            std::printf("x pair %d split over XCDs: My %d nfb %d R %d\n", pair, My, nfb, R);
                            return 1;
                        }
                    }
                }
                cases++;
            }
    std::printf("tile maps: %ld cases ok\n", cases);
    return 0;
}
