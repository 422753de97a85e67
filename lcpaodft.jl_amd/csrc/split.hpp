// split.hpp -- block-dense split-precision forms of the grid contractions.
//
// The home grid is cut into 8x8x8-point blocks.  For each block the sphere
// images (atom, cell) that cover any of its points are listed ("slots") and
// their orbitals concatenated into R rows (padded to a multiple of 64).  The
// block's orbital values are stored once per table set as "tiles" in the
// element type of the form:
//  * bf16x3: hi and lo bf16 planes, each in two fragment orders (below), so
//    that every matrix-core operand is one contiguous 16-byte read per lane;
//  * fp32: [chunk of 32 points][row][32 points, permuted].
// The Hamiltonian of a block is one dense product C_b = Phi_b diag(V) Phi_b^T
// on the matrix cores, kept in 64x64 super-tiles of the upper triangle; every
// computed pair sums its sub-blocks of C_b over the blocks in FP64.
//
// bf16 fragment orders (per block, chunk c of 32 points, lane l = 16 q + m):
//  * row fragments  [c][g = row/16][l][8]: Phi[16 g + m][32 c + 8 q + e]
//    (the A/B operand of v_mfma_f32_16x16x32_bf16 for 16 rows x 32 points);
//  * point fragments [c][pt][g = row/32][l][8]: Phi[32 g + 8 q + e][32 c + 16 pt + m]
//    (the B operand for 32 rows x 16 points).
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
    int64_t sp_off;     // offset of the block's slot-pair table (nslot x nslot)
    int32_t rows;       // R (multiple of 64)
    int32_t g1, g2, g3; // first grid point of the block
    int32_t nslot;      // sphere images covering the block
    int32_t row_off;    // offset of the block's row descriptors (R entries)
    int32_t slot_off;   // offset of the block's slot orbital counts (nslot entries)
    int32_t st_off;     // index of the block's first upper super-tile among all blocks' (H partials)
    int32_t rused;      // orbital rows in use (rows rused .. R - 1 are padding)
    int32_t pad;
};

// Folded density block of the pair formed by two slots a <= b of a block:
// W element (orbital i of a, orbital j of b) is Deff[hoff + i*ld + j], or
// Deff[hoff + j*ld + i] when the computed pair is oriented (b, a).  hoff < 0:
// the two spheres form no computed pair.
struct SlotPair {
    int32_t hoff;
    int32_t meta;       // ld | trans << 6
};

struct HItem {
    int32_t block, st0, nst, pad;   // pad: structure index in a batched launch
};

// Per-structure arguments of the bf16x3 block kernels (by value for one
// structure, as a device array indexed by HItem::pad for a batch).
struct BlockArgs {
    const BlockDev *blocks;
    const uint16_t *hi, *lo;       // row fragments (H)
    const uint16_t *thi, *tlo;     // point fragments (density)
    const int16_t *rowinfo;
    const SlotPair *spair;
    const double *V;               // H: potential of the channel
    const float *Deff;             // density: folded pair blocks (FP32)
    float *part;                   // H: super-tile partials
    double *rho;                   // density: output
    const int16_t *slotno;         // orbital count of each slot
    const float *t32;              // FP32 tiles (the FP32 forms)
    int32_t n1, n2, n3, pad;
};

// items of one row count R, launched together (LDS sized for R)
struct ItemClass {
    int32_t rows, first, count;
};

// H partials: every block's upper 64x64 super-tiles of C_b as the matrix
// cores leave them, super-tile st of block b at float offset
// (st_off(b) + st) * 4096; inside it, 16x16 tile t = 4 rt + ct at t * 256,
// lane l = 16 (r / 4) + c holding rows 4 (l / 16) .. + 3 of column l % 16 as
// four consecutive floats (r, c: row and column inside the tile).
__host__ __device__ __forceinline__ int64_t part_index(int64_t base, int S, int r, int c) {
    const int si = r >> 6, sj = c >> 6;
    const int st = si * S - (si * (si - 1)) / 2 + (sj - si);
    return base + (int64_t)st * 4096 + ((((r & 63) >> 4) * 4 + ((c & 63) >> 4)) * 64 + ((r & 15) >> 2) * 16 + (c & 15)) * 4 +
           (r & 3);
}

// one (pair, block) sub-block of C_b: slots x <= y of the block, whose rows
// start at rx and ry; element (i, j) of H[pair] is C_b[rx + i][ry + j], or
// C_b[rx + j][ry + i] when trans != 0 (the pair's atom A is slot y); an
// element below the diagonal (on-site blocks, x == y) is read from its
// mirror.  base = st_off(block) * 4096, S = super-tiles per block row.
struct Contrib {
    int64_t base;
    int16_t rx, ry, S, trans;
};

struct BlockPlan {
    int nb[3] = {0, 0, 0};
    int64_t nblocks = 0;
    std::vector<BlockDev> blocks;
    std::vector<HItem> items;        // sorted by rows
    std::vector<ItemClass> classes;
    int64_t tile_elems = 0;    // total tile elements (per hi/lo plane)
    int64_t part_tiles = 0;    // total upper super-tiles
    int32_t max_rows = 0;
    int32_t max_slots = 0;
    int32_t max_sub_rows = 0;  // most orbital rows present on one 4x4x4 sub-block (FP64 block density)
    // sphere entries flattened: per (block, sphere)
    int64_t nsph = 0;
    // device
    BlockDev *d_blocks = nullptr;
    HItem *d_items = nullptr;
    // the density split by block size: the super-tiles of the blocks of at
    // most RH_ROWS rows in items of at most RH_WAVES (the 8-wave kernel, two
    // workgroups per CU), split evenly per block, XCD-dealt; the items of
    // the larger blocks (the 16-wave kernel)
    std::vector<HItem> items8, items_big;
    HItem *d_items8 = nullptr, *d_items_big = nullptr;
    int32_t max_rows_big = 0;
    int32_t *d_sph_block = nullptr, *d_sph_atom = nullptr, *d_sph_row0 = nullptr;
    int32_t *d_sph_map = nullptr;  // nsph x BLK_PTS table columns (-1: not covered)
    Contrib *d_contrib = nullptr;
    int32_t *d_contrib_off = nullptr;  // per computed pair in the reduction's launch order (ncanon + 1)
    int32_t *d_red_canon = nullptr;    // the computed pairs in the reduction's launch order
    // tiles of the precision forms
    uint16_t *d_hi = nullptr, *d_lo = nullptr;    // bf16 planes, row fragments
    uint16_t *d_thi = nullptr, *d_tlo = nullptr;  // bf16 planes, point fragments
    int16_t *d_rowinfo = nullptr;                 // per block row: slot * 32 + orbital, -1 for padding
    SlotPair *d_spair = nullptr;
    int32_t *d_contrib_pair = nullptr;
    int64_t ncontrib = 0;
    int32_t red_threads = 64;      // reduce width: the largest hblk_reduce_width of a computed pair
    std::vector<int32_t> red_count;  // contributions per reduction launch slot (host; batches order by it)
    float *d_t32 = nullptr;                     // fp32 tiles (permuted point order)
    bool bf16_ready = false, f32_ready = false;
    float *d_part = nullptr;       // H partials: part_tiles super-tiles of 4096 floats
    int64_t part_elems = 0;
    int16_t *d_slotno = nullptr;   // per block slot
    // bf16x3 Force3: the ordered pair of every slot pair (hoff of D(pair),
    // meta = orbitals of the second slot; hoff < 0: no pair), items over all
    // S x S super-tiles, FP32 gradient tiles, per-(item, slot) sums and the
    // per-atom lists of those sums (item order: deterministic)
    SlotPair *d_fpair = nullptr;
    std::vector<HItem> fitems;
    HItem *d_fitems = nullptr;
    int32_t *d_fbase = nullptr;    // per force item: its first (item, slot) output
    int64_t fout = 0;
    double *d_fout = nullptr;      // fout x 3
    int32_t *d_fatom_off = nullptr, *d_fatom_idx = nullptr;
    float *d_dtile = nullptr;      // 3 x tile_elems floats, gradient fragment order
    bool dtile_ready = false;
};

int plan_build(lcpao_ctx *ctx);
// bf16x3 block kernels, shared with the batched form
BlockArgs block_args(const lcpao_ctx *ctx, const double *V, double *rho);
size_t block_kernel_lds(bool density, int rows, int max_slots, bool persistent, bool db = false);
// whether the double-buffered 16-wave density kernel serves blocks of up to
// `rows` rows and max_slots sphere images (LDS and register limits)
bool rho_db_fits(int rows, int max_slots);
void launch_fill_f64(hipStream_t stream, int64_t n, double *x, double v);
void launch_fill_i32(hipStream_t stream, int64_t n, int32_t *x, int32_t v);
void launch_hblk_reduce(const lcpao_ctx *ctx, hipStream_t stream, double *H);
// args != nullptr: batched launch (item.pad indexes args, `one` unused);
// grid == 0 with a queue: persistent grid (resident workgroups); db: the
// double-buffered density kernel (one workgroup per item; see rho_db_fits)
hipError_t launch_block_kernel(bool density, hipStream_t stream, int rows, int max_slots, const HItem *items,
                               int32_t nitems, const BlockArgs *args, const BlockArgs &one, int32_t *queue,
                               unsigned grid, bool db = false);
// FP32 block kernels (H on the FP32 matrix cores, and the FP32 block
// density): args != nullptr launches the merged items of a batch (item.pad
// indexes args), else one structure (`one`); one workgroup per item
hipError_t launch_f32_block_kernel(bool density, hipStream_t stream, int rows, int max_slots, const HItem *items,
                                   int32_t nitems, const BlockArgs *args, const BlockArgs &one);
// bf16x3 density in items of at most 8 super-tiles, two workgroups per CU
// (blocks of at most 384 rows)
bool rho_w8_fits(int rows);
hipError_t launch_rblk_w8(hipStream_t stream, int rows, const HItem *items, int32_t nitems, const BlockArgs *args,
                          const BlockArgs &one);
// batched fills and reductions of a batch (one launch each for all structures)
struct BatchRed {
    const int32_t *canon;   // computed pairs of the structure, in the reduction's launch order
    const int32_t *coff;    // contribution offsets per launch slot
    const Contrib *contrib;
    const PairDev *pairs;
    const float *part;
    double *H;
    double gridvol;
    int32_t ncanon, pad;
};
struct BatchFill {
    double *x;
    int64_t n;
};
struct BatchFold {
    const int32_t *canon;
    const PairDev *pairs;
    const double *D;        // the structure's flat DM
    float *Deff;            // its FP32 fold
    int32_t ncanon, pad;
};
// the folds and reductions of a batch run over the computed pairs of all
// structures in one flat grid: workgroup w is pair w - off[s] of the
// structure s with off[s] <= w < off[s + 1] (off: nstruct + 1 prefix sums);
// the reduction's workgroup i takes flat pair perm[i] (a permutation of the
// flat pairs: those with the most contributions, over all structures, first)
void launch_deff_f32_batch(hipStream_t stream, int nstruct, const BatchFold *d_fold, const int32_t *d_off,
                           int total_canon);
void launch_hblk_reduce_batch(hipStream_t stream, int nstruct, const BatchRed *d_red, const int32_t *d_off,
                              int total_canon, int threads, const int32_t *d_perm);
// the structure of flat-grid workgroup w (off: nstruct + 1 prefix sums)
__device__ __forceinline__ int batch_struct_of(const int32_t *__restrict__ off, int nstruct, int w) {
    int lo = 0, hi = nstruct;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (__builtin_amdgcn_readfirstlane(off[mid]) <= w) lo = mid;
        else hi = mid;
    }
    return lo;
}
void launch_fill_batch(hipStream_t stream, int nstruct, const BatchFill *d_fill, int64_t max_n);
void plan_free(BlockPlan *p);
int split_tables(lcpao_ctx *ctx, int precision);
// FP32 density in the block form (needs split_tables(ctx, LCPAO_FP32));
// enqueues the fold, the clear and one launch per spin, no synchronisation
int split_density_f32_enqueue(lcpao_ctx *ctx, int nspin, const double *const *d_DM, int64_t dm_offset,
                              double *const *d_rho);
int split_hamiltonian(lcpao_ctx *ctx, int nch, const double *const *d_V, double *const *d_H, int precision);
// FP64 density in the block form (needs the block plan, plan_build); mode 0:
// exact vector FMAs with the orbital columns of each sub-block staged in LDS
// (k_rblk64), 1: the same reading the columns from the tables (k_rblk64t),
// 2: the stacked W tiles of each sub-block on the FP64 matrix cores
// (k_rblk64m); enqueues the fold and one launch per spin, no synchronisation
int split_density_f64_enqueue(lcpao_ctx *ctx, int nspin, const double *const *d_DM, int64_t dm_offset,
                              double *const *d_rho, int mode);
int split_density(lcpao_ctx *ctx, int nspin, const double *const *d_DM, int64_t dm_offset, double *const *d_rho);
// bf16x3 Force3 of one channel: adds w * S_A into d_fsum (natom x 3, atom-major)
int split_force3(lcpao_ctx *ctx, const double *d_V, const double *d_DM, double w, double *d_fsum);
// LCPAO_OK when the bf16x3 Force3 block form holds the geometry (tiles built),
// LCPAO_ERANGE when it does not
int split_force3_ready(lcpao_ctx *ctx);
int fp32_tables(lcpao_ctx *ctx, bool deriv);
void invalidate_split_tables(lcpao_ctx *ctx);
void launch_to_f32(hipStream_t stream, int64_t n, const double *x, float *y);
// folded pair blocks into ctx->d_deff (sym_self: on-site blocks as D + D^T)
void launch_deff_f64(lcpao_ctx *ctx, const double *D, int sym_self);
void launch_deff_f32(lcpao_ctx *ctx, const double *D, int sym_self);  // into d_deff32

}  // namespace lcpao
