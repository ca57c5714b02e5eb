// tdda_tilemap.hpp -- order of the strided passes' tiles within a job.
//
// Workgroups go to the 8 XCDs round robin and each XCD has its own L2; the
// persistent pass grids are multiples of 8 (or one workgroup per tile), so
// tile r of a job runs on XCD (tstart + r) mod 8.  Tiles that share cache
// lines or Green octant entries are therefore placed 8 apart:
//
//   y pass (tile = row x, column block fb): the column blocks of a row split
//   its 128-byte lines between them (rows are 3 Mz values, rarely a multiple
//   of a line) -- all blocks of one row go to one XCD;
//   x pass (tile = row ky, column block fb): the rows ky and My - ky read the
//   same Green octant rows, the blocks of kz and Mz - kz the same entries,
//   and neighbouring blocks share lines -- all blocks of a mirrored pair of
//   rows go to one XCD.
//
// Rows are taken 8 at a time (one per XCD); the last group may be smaller.
#pragma once
#include <hip/hip_runtime.h>

namespace tdda {

// y pass: tile r of nrows x nfb -> (row o, block fb)
__host__ __device__ inline void ytile_map(int r, int nrows, int nfb, int& o, int& fb)
{
    const int o0 = 8 * (r / (8 * nfb));
    const int m = nrows - o0 < 8 ? nrows - o0 : 8;
    const int rr = r - o0 * nfb;
    const int j = rr / m;
    o = o0 + (rr - j * m);
    fb = j;
}

// x pass: tile r of a group of R right-hand sides of one operator, each
// My x nfb tiles -> (row ky, block fb, right-hand side k).  Rows 1..npair
// pair with My - 1 .. My - npair (npair = (My - 1) / 2); all blocks of a
// pair in all R right-hand sides go to one XCD, the R right-hand sides of a
// (row, block) back to back, so that they read the same Green octant rows
// from that XCD's L2.  The unpaired rows 0 and, for even My, My / 2 come
// last.
__host__ __device__ inline void xtile_map(int r, int My, int nfb, int R, int& ky, int& fb, int& k)
{
    const int npair = (My - 1) / 2, per = 2 * nfb * R, paired = npair * per;
    if (r < paired) {
        const int p0 = 8 * (r / (8 * per));
        const int m = npair - p0 < 8 ? npair - p0 : 8;
        const int rr = r - p0 * per;
        const int j = rr / m;
        const int pair = p0 + (rr - j * m);
        k = j % R;
        const int jj = j / R;
        fb = jj >> 1;
        ky = (jj & 1) ? My - 1 - pair : pair + 1;
    } else {
        const int s = r - paired;
        k = s % R;
        const int ss = s / R;
        const int row = ss >= nfb ? 1 : 0;
        fb = ss - row * nfb;
        ky = row ? My / 2 : 0;
    }
}

}  // namespace tdda
