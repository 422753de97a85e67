// split.hpp -- block-dense split-precision forms of the grid contractions.
//
// The home grid is cut into 8x8x8-point blocks.  For each block the sphere
// images (atom, cell) that cover any of its points are listed and their
// orbitals concatenated into R rows (padded to a multiple of 64).  The
// block's orbital values are stored once per table set as "tiles":
// [chunk of 32 points][row][32 points] in the element type of the form
// (bf16 hi/lo pairs, or fp32).  The Hamiltonian of a block is then one
// dense product C_b = Phi_b diag(V) Phi_b^T on the matrix cores, kept in
// 64x64 super-tiles of the upper triangle; every computed pair sums its
// sub-blocks of C_b over the blocks in FP64.
#pragma once
#include "ctx.hpp"

namespace lcpao {

constexpr int BLK = 8;            // points per block edge
constexpr int BLK_PTS = BLK * BLK * BLK;
constexpr int CHUNK = 32;         // points per LDS chunk
constexpr int NCHUNK = BLK_PTS / CHUNK;
constexpr int ST = 64;            // super-tile edge
constexpr int HB_WAVES = 8;       // waves per H block workgroup
constexpr int ST_PER_WAVE = 2;

struct BlockDev {
    int64_t tile_off;   // element offset of the block's tiles
    int64_t part_off;   // super-tile offset of the block's partial C
    int32_t rows;       // R (multiple of 64)
    int32_t g1, g2, g3; // first grid point of the block
};

struct HItem {
    int32_t block, st0, nst, pad;
};

// one (pair, block) sub-block of C_b: element (i, j) of H[pair] is
// C_b[r0 + i][c0 + j] (or C_b[r0 + j][c0 + i] when trans != 0)
struct Contrib {
    int32_t block, r0, c0, trans;
};

struct BlockPlan {
    int nb[3] = {0, 0, 0};
    int64_t nblocks = 0;
    std::vector<BlockDev> blocks;
    std::vector<HItem> items;
    int64_t tile_elems = 0;    // total tile elements (per hi/lo plane)
    int64_t part_tiles = 0;    // total super-tiles
    int32_t max_rows = 0;
    // sphere entries flattened: per (block, sphere)
    int64_t nsph = 0;
    // device
    BlockDev *d_blocks = nullptr;
    HItem *d_items = nullptr;
    int32_t *d_sph_block = nullptr, *d_sph_atom = nullptr, *d_sph_row0 = nullptr;
    int32_t *d_sph_map = nullptr;  // nsph x BLK_PTS table columns (-1: not covered)
    Contrib *d_contrib = nullptr;
    int32_t *d_contrib_off = nullptr;  // per canonical pair (ncanon + 1)
    // tiles of the precision forms
    uint16_t *d_hi = nullptr, *d_lo = nullptr;  // bf16 planes
    uint16_t *d_thi = nullptr, *d_tlo = nullptr;  // transposed bf16 planes [chunk][point][row]
    float *d_w32 = nullptr;                       // density super-tiles W (fp32)
    Contrib *d_contrib_flat = nullptr;            // contributions with their pair (for W assembly)
    int32_t *d_contrib_pair = nullptr;
    int64_t ncontrib = 0;
    float *d_t32 = nullptr;                     // fp32 tiles (permuted point order)
    bool bf16_ready = false, f32_ready = false;
    float *d_part = nullptr;
};

int plan_build(lcpao_ctx *ctx);
void plan_free(BlockPlan *p);
int split_tables(lcpao_ctx *ctx, int precision);
int split_hamiltonian(lcpao_ctx *ctx, int nch, const double *const *d_V, double *const *d_H, int precision);
int split_density(lcpao_ctx *ctx, int nspin, const double *const *d_DM, int64_t dm_offset, double *const *d_rho);
int fp32_tables(lcpao_ctx *ctx, bool deriv);
void invalidate_split_tables(lcpao_ctx *ctx);
void launch_to_f32(hipStream_t stream, int64_t n, const double *x, float *y);
void launch_deff_f64(lcpao_ctx *ctx, const double *D);  // folded pair blocks into ctx->d_deff

}  // namespace lcpao
