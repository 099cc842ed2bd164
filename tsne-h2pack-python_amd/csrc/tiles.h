// tiles.h -- row tiles of the joint P for the attractive pass (see tiles.hip).
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

namespace tsne {

// Largest number of distinct neighbours a tile may have to take the LDS path: one
// gather record per neighbour in LDS (8 B in 2D, 16 B in 3D) within lds_bytes, and
// 16-bit local indices.
// Dynamic LDS available to the staged records: the 160 KiB of a gfx950 CU less what the
// tiled kernel declares statically (its per-group row sums, under 2 KiB).
constexpr int kTileLdsMaxKB = 156;
inline int tile_ucap(int D, int lds_bytes)
{
    const int u = lds_bytes / (D == 2 ? 8 : 16);
    return u < 65536 ? u : 65536;
}

// P's rows split into tiles of `rows` consecutive rows.  Tile b lists the distinct
// columns its rows reference, sorted: cols[off[b] .. off[b+1]); every nonzero of a tile
// with at most tile_ucap(D) distinct columns has its column replaced by the position
// of that column in the tile's list (lidx, 16 bits).  Valid while P's numbering stays.
struct AttrTiles {
    int rows = 0, ntiles = 0;
    int umax = 0;          // largest distinct count among the LDS-path tiles
    int n_global = 0;      // tiles above the LDS limit (they gather from HBM)
    long long ucount = 0;  // sum of the distinct counts
    int *off = nullptr;    // [ntiles + 1]
    int *cols = nullptr;   // [ucount]
    uint16_t *lidx = nullptr;  // [nnz + 8]
};

// Build the tiles of the CSR (rp, col) with n rows and nnz nonzeros (columns sorted or
// not).  Temporaries are stream-ordered allocations released before returning.
int tiles_build(AttrTiles &t, int n, int nnz, const int *rp, const int *col, int rows, int D,
                int ucap, hipStream_t st);
void tiles_free(AttrTiles &t, hipStream_t st);

}  // namespace tsne
