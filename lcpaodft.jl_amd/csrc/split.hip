// split.hip -- block-dense split-precision Hamiltonian on the gfx950 matrix
// cores, plus the FP32 table copies used by the split-precision density and
// Force3.  See split.hpp for the data layout.
#include <algorithm>
#include <cstring>
#include <thread>
#include <unordered_map>

#include "split.hpp"

namespace lcpao {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));

#define SPLIT_TRY(ctx, call)                                                        \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                     \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);        \
            return LCPAO_EDEVICE;                                                   \
        }                                                                           \
    } while (0)

constexpr int MAX_ROWS = 576;    // R limit (bf16x3 H: four R x 32 bf16 planes + V in 160 KB of LDS)
constexpr int LROW32 = 36;       // LDS row stride of fp32 tiles (32 + 4 pad)
constexpr int FB_WAVES = 8;      // waves of the Force3 workgroup, two super-tiles of one row each
constexpr int MAX_SLOT_PAIRS = 2048;  // slot pairs of a block (tables staged in LDS by the block kernels)
constexpr int RH_WAVES = 8;           // waves of the two-per-CU density workgroup, one super-tile each
constexpr int RH_ROWS = 384;          // the largest block that workgroup holds

__device__ __forceinline__ int st_index(int si, int sj, int S) { return si * S - (si * (si - 1)) / 2 + (sj - si); }

// position of point q (0..31) inside a fp32 chunk row: lane group q%4 of the
// 16x16x4 layout reads its 8 points contiguously
__host__ __device__ __forceinline__ int perm32(int q) { return (q & 3) * 8 + (q >> 2); }

// element offset, inside a block's tiles, of (row, point 32 c + k) in the two
// bf16 fragment orders
__host__ __device__ __forceinline__ int64_t rowfrag_index(int c, int row, int k, int R) {
    return ((((int64_t)c * (R / 16) + row / 16) * 64 + (k / 8) * 16 + row % 16) * 8) + k % 8;
}
__host__ __device__ __forceinline__ int64_t ptfrag_index(int c, int row, int k, int R) {
    const int rr = row % 32;
    return ((((int64_t)c * 2 + k / 16) * (R / 32) + row / 32) * 64 + (rr / 8) * 16 + k % 16) * 8 + rr % 8;
}

// ----------------------------------------------------------------------------
// tile construction (once per table set)
// ----------------------------------------------------------------------------
__global__ void k_tiles(int64_t nsph, const int32_t *__restrict__ sph_block, const int32_t *__restrict__ sph_atom,
                        const int32_t *__restrict__ sph_row0, const int32_t *__restrict__ sph_map,
                        const BlockDev *__restrict__ blocks, const int64_t *__restrict__ tab_off,
                        const int32_t *__restrict__ no, const double *__restrict__ tab, int mode,
                        uint16_t *__restrict__ hi, uint16_t *__restrict__ lo, float *__restrict__ t32,
                        uint16_t *__restrict__ thi, uint16_t *__restrict__ tlo) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nsph * BLK_PTS) return;
    const int64_t s = gid / BLK_PTS;
    const int k = (int)(gid - s * BLK_PTS);
    const int t = sph_map[gid];
    const int a = sph_atom[s];
    const int NO = no[a];
    const BlockDev B = blocks[sph_block[s]];
    const int c = k / CHUNK, q = k % CHUNK;
    const double *col = tab + tab_off[a] + (int64_t)(t < 0 ? 0 : t) * NO;
    for (int i = 0; i < NO; i++) {
        const double x = t >= 0 ? col[i] : 0.0;
        const int64_t row = sph_row0[s] + i;
        if (mode == LCPAO_BF16X3) {
            const float xf = (float)x;
            const __bf16 h = (__bf16)xf;
            const __bf16 l = (__bf16)(xf - (float)h);
            const int64_t idx = B.tile_off + rowfrag_index(c, (int)row, q, B.rows);
            hi[idx] = __builtin_bit_cast(uint16_t, h);
            lo[idx] = __builtin_bit_cast(uint16_t, l);
            const int64_t tdx = B.tile_off + ptfrag_index(c, (int)row, q, B.rows);
            thi[tdx] = __builtin_bit_cast(uint16_t, h);
            tlo[tdx] = __builtin_bit_cast(uint16_t, l);
        } else {
            const int64_t idx = B.tile_off + ((int64_t)c * B.rows + row) * CHUNK + perm32(q);
            t32[idx] = (float)x;
        }
    }
}

__global__ void k_to_f32(int64_t n, const double *__restrict__ x, float *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (float)x[i];
}

void launch_to_f32(hipStream_t stream, int64_t n, const double *x, float *y) {
    if (n > 0) hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, n, x, y);
}

// ----------------------------------------------------------------------------
// block Hamiltonian: C_b = Phi diag(V) Phi^T, upper 64x64 super-tiles
// ----------------------------------------------------------------------------
__device__ __forceinline__ void st_decode(int st, int S, int &si, int &sj) {
    si = 0;
    while (st >= S - si) { st -= S - si; si++; }
    sj = si + st;
}

// Pointers that a kernel reads from device memory (the batched launches'
// per-structure arguments) are generic to the compiler, which then addresses
// them with 64-bit flat VGPR addresses; reading through global-address-space
// pointers lets it use SGPR bases and global loads, as for kernel arguments.
// (The host pass parses these device functions too, and its targets have no
// such address space: there the qualifier is empty.)
#if defined(__HIP_DEVICE_COMPILE__)
#define GLOBAL_AS __attribute__((address_space(1)))
#else
#define GLOBAL_AS
#endif
template <class T>
__device__ __forceinline__ GLOBAL_AS T *as_global(T *p) {
    return (GLOBAL_AS T *)p;
}

__device__ __forceinline__ float bf_bits_to_f32(uint16_t u) { return __uint_as_float((uint32_t)u << 16); }

// ---- bf16x3 form ------------------------------------------------------------
// One workgroup per item (a block and up to 16 of its super-tiles, two per
// wave, accumulators in registers over all 512 points).  Per 32-point chunk
// the LDS holds four row-fragment planes: Phi hi, Phi lo, (V Phi) hi and
// (V Phi) lo, each R x 32 bf16; V Phi is formed once per element while the
// chunk is staged.  Every operand read is one ds_read_b128 at lane * 16 bytes
// of a 1 KB fragment.  The next chunk's fragments are loaded into registers
// before the current chunk is multiplied, so HBM latency overlaps the MFMAs.
constexpr int FRAG = 512;                          // bf16 elements of one fragment
constexpr int MAXF = MAX_ROWS / 16 / HB_WAVES;     // fragments per wave and plane (R = MAX_ROWS)
// NF: fragments per wave and plane for the item class, ceil(R / 16 / HB_WAVES)

// L2 warm-up: one dword of every 128-byte line of a chunk of a block's hi and
// lo planes (32 R bf16 each, R / 2 lines per plane).  The values are kept
// until the chunk is staged (by then the loads have completed in order) and
// only then released, so the touch never stalls the wave.
struct Touch {
    uint32_t a, b;
};
__device__ __forceinline__ Touch touch_chunk(const GLOBAL_AS uint16_t *__restrict__ hi,
                                             const GLOBAL_AS uint16_t *__restrict__ lo, int64_t off, int R, int tid) {
    // always exactly two loads per thread (indices past the chunk repeat line
    // 0), so that the compiler can count them when it places its waits
    Touch t;
    const int half = R / 2;
    int x = tid < R ? tid : 0;
    t.a = *reinterpret_cast<const GLOBAL_AS uint32_t *>((x < half ? hi : lo) + off + (int64_t)(x < half ? x : x - half) * 64);
    x = tid + 64 * HB_WAVES < R ? tid + 64 * HB_WAVES : 0;
    t.b = *reinterpret_cast<const GLOBAL_AS uint32_t *>((x < half ? hi : lo) + off + (int64_t)(x < half ? x : x - half) * 64);
    return t;
}
__device__ __forceinline__ void release(const Touch &t) { asm volatile("" ::"v"(t.a), "v"(t.b)); }

// L2 warm-up of a chunk by NT threads: one dword per 128-byte line, K loads
// per thread (lines past the chunk repeat line 0), as touch_chunk.
template <int NT>
struct TouchN {
    static constexpr int K = (MAX_ROWS + NT - 1) / NT;
    uint32_t v[K];
};
template <int NT>
__device__ __forceinline__ TouchN<NT> touch_chunk_n(const GLOBAL_AS uint16_t *__restrict__ hi,
                                                    const GLOBAL_AS uint16_t *__restrict__ lo, int64_t off, int R,
                                                    int tid) {
    TouchN<NT> t;
    const int half = R / 2;
#pragma unroll
    for (int k = 0; k < TouchN<NT>::K; k++) {
        const int x = tid + k * NT < R ? tid + k * NT : 0;
        t.v[k] = *reinterpret_cast<const GLOBAL_AS uint32_t *>((x < half ? hi : lo) + off +
                                                               (int64_t)(x < half ? x : x - half) * 64);
    }
    return t;
}
template <int NT>
__device__ __forceinline__ void release_n(const TouchN<NT> &t) {
#pragma unroll
    for (int k = 0; k < TouchN<NT>::K; k++) asm volatile("" ::"v"(t.v[k]));
}

// Work distribution of the block kernels (MODE, a template argument, so that
// each schedule compiles to a kernel of its own and the one-pass forms carry
// no loop).  LAUNCH_ONE: one structure (`one`), one workgroup per item (grid =
// item count).  LAUNCH_BATCH: item.pad selects the structure's entry of
// `args`, one workgroup per item.  LAUNCH_QUEUE: as LAUNCH_BATCH, but the grid
// is persistent and each workgroup takes the next item index from the queue
// counter (zeroed before the launch) until the list is exhausted, so every
// workgroup reaches the exit.
enum { LAUNCH_ONE = 0, LAUNCH_BATCH = 1, LAUNCH_QUEUE = 2 };
template <int MODE, class F>
__device__ __forceinline__ void run_items(const HItem *__restrict__ items, int32_t nitems,
                                          const BlockArgs *__restrict__ args, const BlockArgs &one,
                                          int32_t *__restrict__ queue, F body) {
    if constexpr (MODE == LAUNCH_ONE) {
        if ((int)blockIdx.x < nitems) body(items[blockIdx.x], one);
    } else if constexpr (MODE == LAUNCH_BATCH) {
        if ((int)blockIdx.x < nitems) {
            const HItem it = items[blockIdx.x];
            body(it, args[it.pad]);
        }
    } else {
        __shared__ int32_t s_next;
        for (;;) {
            __syncthreads();  // the previous item is done with the LDS and s_next
            if (threadIdx.x == 0) s_next = atomicAdd(queue, 1);
            __syncthreads();
            const int32_t i = s_next;
            if (i >= nitems) return;
            const HItem it = items[i];
            body(it, args[it.pad]);
        }
    }
}
inline int launch_mode(const BlockArgs *args, const int32_t *queue) {
    return args == nullptr ? LAUNCH_ONE : queue == nullptr ? LAUNCH_BATCH : LAUNCH_QUEUE;
}

// Writes a wave's super-tile accumulators as they stand (the H partial layout
// of part_index): tile t = 4 rt + ct of global super-tile st at
// part[(st * 16 + t) * 256 + lane * 4], one 16-byte store per lane, so each
// store instruction writes 1 KB of consecutive bytes.  Tiles that no pair
// reads are skipped: those below the diagonal of a diagonal super-tile (the
// reduction reads their mirror) and those whose rows or columns are all
// padding.  The conditions are uniform across the wave.
__device__ __forceinline__ void store_tiles(const f32x4 *acc, int si, int sj, int64_t st, int rused, int lane,
                                            GLOBAL_AS float *__restrict__ part) {
    GLOBAL_AS f32x4 *dst = reinterpret_cast<GLOBAL_AS f32x4 *>(part) + st * 16 * 64 + lane;
#pragma unroll
    for (int rt = 0; rt < 4; rt++)
#pragma unroll
        for (int ct = 0; ct < 4; ct++) {
            if (si == sj && rt > ct) continue;
            if (si * ST + rt * 16 >= rused || sj * ST + ct * 16 >= rused) continue;
            dst[(rt * 4 + ct) * 64] = acc[rt * 4 + ct];
        }
}

inline int nfrag(int rows) { return (rows / 16 + HB_WAVES - 1) / HB_WAVES; }

__device__ __forceinline__ void tile_bf16x3(const u16x8 *__restrict__ PH, const u16x8 *__restrict__ PL,
                                            const u16x8 *__restrict__ VH, const u16x8 *__restrict__ VL, int si,
                                            int sj, int lane, f32x4 *acc) {
    bf16x8 ah[4], al[4];
#pragma unroll
    for (int rt = 0; rt < 4; rt++) {
        ah[rt] = __builtin_bit_cast(bf16x8, PH[(si * 4 + rt) * 64 + lane]);
        al[rt] = __builtin_bit_cast(bf16x8, PL[(si * 4 + rt) * 64 + lane]);
    }
#pragma unroll
    for (int ct = 0; ct < 4; ct++) {
        const bf16x8 bh = __builtin_bit_cast(bf16x8, VH[(sj * 4 + ct) * 64 + lane]);
        const bf16x8 bl = __builtin_bit_cast(bf16x8, VL[(sj * 4 + ct) * 64 + lane]);
#pragma unroll
        for (int rt = 0; rt < 4; rt++) {
            f32x4 c = acc[rt * 4 + ct];
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[rt], bh, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rt], bl, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rt], bh, c, 0, 0, 0);
            acc[rt * 4 + ct] = c;
        }
    }
}


template <int NF>
__device__ __forceinline__ void hblk_item(const HItem it, const BlockArgs &A, unsigned char *lds_raw) {
    const BlockDev B = as_global(A.blocks)[it.block];
    const GLOBAL_AS uint16_t *__restrict__ hi = as_global(A.hi);
    const GLOBAL_AS uint16_t *__restrict__ lo = as_global(A.lo);
    const GLOBAL_AS double *__restrict__ V = as_global(A.V);
    GLOBAL_AS float *__restrict__ part = as_global(A.part);
    const int n1 = A.n1, n2 = A.n2, n3 = A.n3;
    const int R = B.rows, S = R / ST, G = R / 16;
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63, q = lane >> 4;
    u16x8 *PH = reinterpret_cast<u16x8 *>(lds_raw);
    u16x8 *PL = PH + G * 64;
    u16x8 *VH = PL + G * 64;
    u16x8 *VL = VH + G * 64;
    float *Vb = reinterpret_cast<float *>(VL + G * 64);  // V at the block's 512 points
    {
        const int k = tid;  // 512 threads, one point each
        const int x1 = B.g1 + k / (BLK * BLK), x2 = B.g2 + (k / BLK) % BLK, x3 = B.g3 + k % BLK;
        Vb[k] = (x1 < n1 && x2 < n2 && x3 < n3) ? (float)V[((int64_t)x1 * n2 + x2) * n3 + x3] : 0.f;
    }
    const bool has0 = wave < it.nst, has1 = wave + HB_WAVES < it.nst;
    int si0 = 0, sj0 = 0, si1 = 0, sj1 = 0;
    if (has0) st_decode(it.st0 + wave, S, si0, sj0);
    if (has1) st_decode(it.st0 + wave + HB_WAVES, S, si1, sj1);
    f32x4 acc0[16], acc1[16];
#pragma unroll
    for (int t = 0; t < 16; t++) {
        acc0[t] = f32x4{0.f, 0.f, 0.f, 0.f};
        acc1[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    u16x8 fh[NF], fl[NF];
    // chunk c: fragments g = wave + 8 f of both planes
    auto prefetch = [&](int c) {
        const GLOBAL_AS uint16_t *sh = hi + B.tile_off + (int64_t)c * G * FRAG;
        const GLOBAL_AS uint16_t *sl = lo + B.tile_off + (int64_t)c * G * FRAG;
#pragma unroll
        for (int f = 0; f < NF; f++) {  // unconditional (clamped) loads: see touch_chunk
            const int g = min(wave + f * HB_WAVES, G - 1);
            fh[f] = *reinterpret_cast<const GLOBAL_AS u16x8 *>(sh + (g * 64 + lane) * 8);
            fl[f] = *reinterpret_cast<const GLOBAL_AS u16x8 *>(sl + (g * 64 + lane) * 8);
        }
    };
    // chunk c is touched two iterations before it is staged (c - 2), loaded
    // into registers one iteration before (c - 1); touches alternate between
    // two register slots so that no register copy waits on a load in flight
    auto step = [&](int c, Touch &slot) {
        __syncthreads();
        // V at this lane's eight points k = 32 c + 8 q + e
        const f32x8 vv = *reinterpret_cast<const f32x8 *>(&Vb[c * CHUNK + 8 * q]);
        // branch-free: a wave past the last fragment rewrites fragment G - 1
        // with the same values (no branch keeps the compiler's load counts exact)
#pragma unroll
        for (int f = 0; f < NF; f++) {
            const int g = min(wave + f * HB_WAVES, G - 1);
            const int at = g * 64 + lane;
            PH[at] = fh[f];
            PL[at] = fl[f];
            u16x8 xh, xl;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const float x = (bf_bits_to_f32(fh[f][e]) + bf_bits_to_f32(fl[f][e])) * vv[e];
                const __bf16 h = (__bf16)x;
                xh[e] = __builtin_bit_cast(uint16_t, h);
                xl[e] = __builtin_bit_cast(uint16_t, (__bf16)(x - (float)h));
            }
            VH[at] = xh;
            VL[at] = xl;
        }
        release(slot);
        __syncthreads();
        // c is a constant in the unrolled loop: these branches fold away
        if (c + 1 < NCHUNK) prefetch(c + 1);
        if (c + 2 < NCHUNK) slot = touch_chunk(hi, lo, B.tile_off + (int64_t)(c + 2) * G * FRAG, R, tid);
        if (has0) tile_bf16x3(PH, PL, VH, VL, si0, sj0, lane, acc0);
        if (has1) tile_bf16x3(PH, PL, VH, VL, si1, sj1, lane, acc1);
    };
    Touch t0{0u, 0u}, t1 = touch_chunk(hi, lo, B.tile_off + (int64_t)G * FRAG, R, tid);
    prefetch(0);
    // fully unrolled: with no loop back-edge the compiler's waits count the
    // loads exactly and never wait for a touch issued in the step before
#pragma unroll
    for (int c = 0; c < NCHUNK; c += 2) {
        step(c, t0);
        step(c + 1, t1);
    }
    // lane holds D[4q + v][l16] of each 16x16 tile
    const int64_t st = (int64_t)B.st_off + it.st0 + wave;
    if (has0) store_tiles(acc0, si0, sj0, st, B.rused, lane, part);
    if (has1) store_tiles(acc1, si1, sj1, st + HB_WAVES, B.rused, lane, part);
}

template <int NF, int MODE>
__global__ __launch_bounds__(64 * HB_WAVES) void k_hblk_bf(const HItem *__restrict__ items, int32_t nitems,
                                                          const BlockArgs *__restrict__ args, BlockArgs one,
                                                          int32_t *__restrict__ queue) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    run_items<MODE>(items, nitems, args, one, queue,
                    [&](const HItem &it, const BlockArgs &A) { hblk_item<NF>(it, A, lds_raw); });
}

// ---- fp32 form --------------------------------------------------------------
// DIAG: a diagonal super-tile (si == sj), whose 16x16 tiles below the
// diagonal (rt > ct) hold only mirrored elements that the reduction never
// reads: their products are skipped (6 of 16 tiles).  (The bf16x3 H kernel
// keeps one path: two inlined versions of its tile made it spill.)
template <bool DIAG>
__device__ __forceinline__ void tile_f32(const float *__restrict__ A32, const float *__restrict__ Vs, int si, int sj,
                                         int l16, int q, f32x4 *acc) {
    f32x8 a[4];
#pragma unroll
    for (int rt = 0; rt < 4; rt++) {
        const int row = si * ST + rt * 16 + l16;
        a[rt] = *reinterpret_cast<const f32x8 *>(&A32[row * LROW32 + q * 8]);
    }
#pragma unroll
    for (int ct = 0; ct < 4; ct++) {
        const int row = sj * ST + ct * 16 + l16;
        f32x8 b = *reinterpret_cast<const f32x8 *>(&A32[row * LROW32 + q * 8]);
#pragma unroll
        for (int e = 0; e < 8; e++) b[e] *= Vs[q * 8 + e];
#pragma unroll
        for (int rt = 0; rt < 4; rt++) {
            if (DIAG && rt > ct) continue;
            f32x4 c = acc[rt * 4 + ct];
#pragma unroll
            for (int s = 0; s < 8; s++) c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rt][s], b[s], c, 0, 0, 0);
            acc[rt * 4 + ct] = c;
        }
    }
}
__device__ __forceinline__ void tile_f32_any(const float *__restrict__ A32, const float *__restrict__ Vs, int si, int sj,
                                             int l16, int q, f32x4 *acc) {
    if (si == sj) tile_f32<true>(A32, Vs, si, sj, l16, q, acc);
    else tile_f32<false>(A32, Vs, si, sj, l16, q, acc);
}

__device__ __forceinline__ void hblk_f32_item(const HItem it, const BlockArgs &A, unsigned char *lds_raw) {
    const BlockDev B = as_global(A.blocks)[it.block];
    const GLOBAL_AS float *__restrict__ t32 = as_global(A.t32);
    const GLOBAL_AS double *__restrict__ V = as_global(A.V);
    GLOBAL_AS float *__restrict__ part = as_global(A.part);
    const int n1 = A.n1, n2 = A.n2, n3 = A.n3;
    const int R = B.rows, S = R / ST;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, l16 = lane & 15, q = lane >> 4;
    float *A32 = reinterpret_cast<float *>(lds_raw);
    float *Vs = A32 + R * LROW32;
    const bool has0 = wave < it.nst, has1 = wave + HB_WAVES < it.nst;
    int si0 = 0, sj0 = 0, si1 = 0, sj1 = 0;
    if (has0) st_decode(it.st0 + wave, S, si0, sj0);
    if (has1) st_decode(it.st0 + wave + HB_WAVES, S, si1, sj1);
    f32x4 acc0[16], acc1[16];
#pragma unroll
    for (int t = 0; t < 16; t++) {
        acc0[t] = f32x4{0.f, 0.f, 0.f, 0.f};
        acc1[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    for (int c = 0; c < NCHUNK; c++) {
        __syncthreads();
        const GLOBAL_AS float *s32 = t32 + B.tile_off + (int64_t)c * R * CHUNK;
        for (int e = tid * 4; e < R * CHUNK; e += 64 * HB_WAVES * 4) {
            const int row = e / CHUNK, col = e % CHUNK;
            *reinterpret_cast<f32x4 *>(&A32[row * LROW32 + col]) = *reinterpret_cast<const f32x4 *>(&s32[e]);
        }
        if (tid < CHUNK) {
            const int k = c * CHUNK + tid;
            const int x1 = B.g1 + k / (BLK * BLK), x2 = B.g2 + (k / BLK) % BLK, x3 = B.g3 + k % BLK;
            const float v = (x1 < n1 && x2 < n2 && x3 < n3) ? (float)V[((int64_t)x1 * n2 + x2) * n3 + x3] : 0.f;
            Vs[perm32(tid)] = v;
        }
        __syncthreads();
        if (has0) tile_f32_any(A32, Vs, si0, sj0, l16, q, acc0);
        if (has1) tile_f32_any(A32, Vs, si1, sj1, l16, q, acc1);
    }
    const int64_t st = (int64_t)B.st_off + it.st0 + wave;
    if (has0) store_tiles(acc0, si0, sj0, st, B.rused, lane, part);
    if (has1) store_tiles(acc1, si1, sj1, st + HB_WAVES, B.rused, lane, part);
}

template <int MODE>
__global__ __launch_bounds__(64 * HB_WAVES) void k_hblk_f32(const HItem *__restrict__ items, int32_t nitems,
                                                           const BlockArgs *__restrict__ args, BlockArgs one) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    run_items<MODE>(items, nitems, args, one, nullptr,
                    [&](const HItem &it, const BlockArgs &A) { hblk_f32_item(it, A, lds_raw); });
}

// per computed pair: FP64 sum of its sub-blocks of C_b over the blocks, x
// GridVol; writes H[P] and H[P']^T.  One workgroup per pair.  A sub-block is
// read as the 16-byte lane slots of the tile layout (part_index: four
// consecutive rows of one column): thread t takes column b = t mod ncol of
// the sub-block and row group g = t / ncol of the 4-row groups its rows span,
// so the lanes of a wave read consecutive slots.  Value q of a slot is tile
// row 4 g + q of the span, i.e. sub-block row 4 g + q - (rx mod 4); the
// thread's four accumulators therefore keep one meaning while the class
// (trans, rx mod 4) stays the same.  A pair's contributions are sorted by
// class (then by block), and at each change of class the accumulators are
// added into the pair's block in LDS (each element by one thread, a barrier
// between classes), so the sum order is fixed.  An on-site pair (slot x =
// slot y) reads the diagonal and the upper triangle, which the tiles hold,
// and adds each value above the diagonal to its mirror element as well.
constexpr int RED_BATCH = 256;
constexpr int RED_UNROLL = 8;
constexpr int RED_MAXNO = 32;  // the split-precision plan holds at most 32 orbitals per atom
// workgroup width: both orientations' column x row-group grids of the pair
__host__ __device__ inline int hblk_reduce_width(int noa, int nob) {
    const int w = max(nob * ((noa + 6) / 4), noa * ((nob + 6) / 4));
    return (w + 63) / 64 * 64;
}
__device__ __forceinline__ void hblk_reduce_pair(int64_t ci, const int32_t *__restrict__ canon,
                                                 const int32_t *__restrict__ coff,
                                                 const Contrib *__restrict__ contrib,
                                                 const PairDev *__restrict__ pairs, const float *__restrict__ part,
                                                 double gridvol, double *__restrict__ H) {
    __shared__ int64_t s_base[RED_BATCH];
    __shared__ int32_t s_rxy[RED_BATCH], s_st[RED_BATCH];
    __shared__ double s_h[RED_MAXNO * RED_MAXNO];
    const int p = canon[ci];
    const PairDev P = pairs[p];
    const int noa = P.noa, nob = P.nob, ne = noa * nob;
    if (ne == 0) return;  // an atom without orbitals: nothing to sum (the whole workgroup leaves)
    const int k0 = coff[ci], k1 = coff[ci + 1];
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int e = tid; e < ne; e += nt) s_h[e] = 0.0;
    // (column, row group) of this thread with the pair's rows (trans = 0) or
    // columns (trans = 1) along the tile rows
    const int bN = tid % nob, gN = tid / nob, bT = tid % noa, gT = tid / noa;
    int cls = -1;  // class of the accumulators: trans * 4 + rx mod 4
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    bool onsite = false;
    auto flush = [&]() {
        const int tr = cls >> 2, sh = cls & 3, nrow = tr ? nob : noa;
        const int b = tr ? bT : bN, g = tr ? gT : gN;
        if (g < ((sh + nrow + 3) >> 2)) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int a = 4 * g + q - sh;
                if (a < 0 || a >= nrow) continue;
                if (!onsite) {
                    s_h[tr ? b * nob + a : a * nob + b] += acc[q];
                } else {  // C_b[rx + a][rx + b], a <= b, is element (a, b) and, below the diagonal, (b, a)
                    if (a <= b) s_h[a * nob + b] += acc[q];
                    if (a < b) s_h[b * nob + a] += acc[q];
                }
            }
        }
    };
    for (int kb = k0; kb < k1; kb += RED_BATCH) {
        const int nb = min(RED_BATCH, k1 - kb);
        __syncthreads();
        for (int x = tid; x < nb; x += nt) {  // the width may be below RED_BATCH
            const Contrib C = contrib[kb + x];
            s_base[x] = C.base;
            s_rxy[x] = (int32_t)C.rx | ((int32_t)C.ry << 16);
            s_st[x] = (int32_t)C.S | ((int32_t)C.trans << 16);
        }
        __syncthreads();
        // an on-site pair (atom A with itself, no lattice shift) is the one
        // whose sub-blocks sit on the diagonal (slot x = slot y)
        if (kb == k0) onsite = (s_st[0] >> 16) == 0 && (s_rxy[0] & 0xffff) == (s_rxy[0] >> 16);
        for (int k = 0; k < nb; k += RED_UNROLL) {
            f32x4 v[RED_UNROLL];
#pragma unroll
            for (int u = 0; u < RED_UNROLL; u++) {
                const int kk = min(k + u, nb - 1);
                const int rxy = s_rxy[kk], sst = s_st[kk];
                const int rx = rxy & 0xffff, ry = rxy >> 16, tr = sst >> 16;
                const int nrow = tr ? nob : noa;
                // a thread past the span reads the span's last row group
                const int g = min(tr ? gT : gN, ((rx & 3) + nrow - 1) >> 2);
                const int c = ry + (tr ? bT : bN);
                int r = (rx & ~3) + 4 * g;
                // on-site: rows wholly below the diagonal are not needed (their
                // mirror is), and their tile may not be stored; read the
                // diagonal's row group instead
                if (onsite && r > c) r = c & ~3;
                const int64_t at = part_index(s_base[kk], sst & 0xffff, r, c);
                v[u] = *reinterpret_cast<const f32x4 *>(part + at);
            }
#pragma unroll
            for (int u = 0; u < RED_UNROLL; u++) {
                if (k + u >= nb) break;
                const int c = ((s_st[k + u] >> 16) << 2) | (s_rxy[k + u] & 3);
                if (c != cls) {  // the same for every thread of the workgroup
                    if (cls >= 0) flush();
                    __syncthreads();
                    cls = c;
                    acc[0] = acc[1] = acc[2] = acc[3] = 0.0;
                }
                acc[0] += (double)v[u][0];
                acc[1] += (double)v[u][1];
                acc[2] += (double)v[u][2];
                acc[3] += (double)v[u][3];
            }
        }
    }
    if (cls >= 0) flush();
    __syncthreads();
    for (int e = tid; e < ne; e += nt) {
        const int i = e / nob, j = e - i * nob;
        const double s = s_h[e] * gridvol;
        H[P.hoff + e] = s;
        if (P.partner >= 0 && P.partner != p) H[pairs[P.partner].hoff + (int64_t)j * noa + i] = s;
    }
}

__global__ __launch_bounds__(1024) void k_hblk_reduce(const int32_t *__restrict__ canon, int64_t ncanon,
                                                      const int32_t *__restrict__ coff,
                                                      const Contrib *__restrict__ contrib,
                                                      const PairDev *__restrict__ pairs, const float *__restrict__ part,
                                                      double gridvol, double *__restrict__ H) {
    const int64_t ci = blockIdx.x;
    if (ci >= ncanon) return;
    hblk_reduce_pair(ci, canon, coff, contrib, pairs, part, gridvol, H);
}

// the reductions of every structure of a batch in one launch over the
// computed pairs of all structures (a flat grid: no idle workgroups for the
// smaller structures)
__global__ __launch_bounds__(1024) void k_hblk_reduce_batch(const BatchRed *__restrict__ red,
                                                            const int32_t *__restrict__ off, int nstruct,
                                                            const int32_t *__restrict__ perm) {
    const int w = perm[blockIdx.x];
    const int st = batch_struct_of(off, nstruct, w);
    const BatchRed R = red[st];
    const int64_t ci = w - off[st];
    if (ci >= R.ncanon) return;
    hblk_reduce_pair(ci, R.canon, R.coff, R.contrib, R.pairs, R.part, R.gridvol, R.H);
}

__global__ void k_fill_batch(const BatchFill *__restrict__ fill) {
    const BatchFill F = fill[blockIdx.y];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < F.n) F.x[i] = 0.0;
}

// ----------------------------------------------------------------------------
// host side
// ----------------------------------------------------------------------------
namespace {

template <class T>
int up(lcpao_ctx *ctx, T **dst, const std::vector<T> &v) {
    if (*dst) { (void)hipFree(*dst); *dst = nullptr; }
    SPLIT_TRY(ctx, hipMalloc(reinterpret_cast<void **>(dst), std::max<size_t>(1, v.size()) * sizeof(T)));
    if (!v.empty()) SPLIT_TRY(ctx, hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return LCPAO_OK;
}

void cell_ijk3(int cpy, int cell, int *ijk) {
    if (cell == 0) { ijk[0] = ijk[1] = ijk[2] = 0; return; }
    const int W = 2 * cpy + 1;
    const int zero_lin = (cpy * W + cpy) * W + cpy;
    const int lin = cell <= zero_lin ? cell - 1 : cell;
    ijk[0] = lin / (W * W) - cpy;
    ijk[1] = (lin / W) % W - cpy;
    ijk[2] = lin % W - cpy;
}

// XCD-aware order for the one-workgroup-per-item launches: workgroup i runs
// on XCD i % 8 (round-robin dispatch), so the items are dealt from 8 slabs of
// blocks along n1 -- slab x fills positions x, x + 8, ... -- and within a slab
// they follow the block order.  The items that run together on one XCD are
// then neighbouring blocks whose density-matrix gathers read the same pair
// blocks through that XCD's L2.
void deal_xcd(const BlockPlan *P, std::vector<HItem> &items, int n1) {
    constexpr int NXCD = 8;
    std::vector<std::vector<HItem>> slab(NXCD);
    std::vector<HItem> sorted = items;
    std::stable_sort(sorted.begin(), sorted.end(), [](const HItem &x, const HItem &y) {
        return x.block != y.block ? x.block < y.block : x.st0 < y.st0;
    });
    for (const HItem &it : sorted) {
        const int64_t g1 = P->blocks[it.block].g1;
        slab[std::min<int64_t>(NXCD - 1, g1 * NXCD / std::max(1, n1))].push_back(it);
    }
    size_t k = 0;
    std::vector<size_t> pos(NXCD, 0);
    while (k < items.size())
        for (int x = 0; x < NXCD; x++)
            if (pos[x] < slab[x].size()) items[k++] = slab[x][pos[x]++];
}

}  // namespace

void plan_free(BlockPlan *p) {
    if (!p) return;
    void *ptrs[] = {p->d_blocks, p->d_items, p->d_items8, p->d_items_big, p->d_sph_block, p->d_sph_atom, p->d_sph_row0, p->d_sph_map,
                    p->d_contrib, p->d_contrib_off, p->d_red_canon, p->d_hi, p->d_lo, p->d_t32, p->d_part, p->d_thi, p->d_tlo,
                    p->d_rowinfo, p->d_spair, p->d_contrib_pair, p->d_slotno, p->d_fpair,
                    p->d_fitems, p->d_fbase, p->d_fout, p->d_fatom_off, p->d_fatom_idx, p->d_dtile};
    for (void *x : ptrs)
        if (x) (void)hipFree(x);
    delete p;
}

void invalidate_split_tables(lcpao_ctx *ctx) {
    if (ctx->plan) { ctx->plan->bf16_ready = false; ctx->plan->f32_ready = false; ctx->plan->dtile_ready = false; }
    if (ctx->d_tab32) { (void)hipFree(ctx->d_tab32); ctx->d_tab32 = nullptr; }
    for (int x = 0; x < 3; x++)
        if (ctx->d_dtab32[x]) { (void)hipFree(ctx->d_dtab32[x]); ctx->d_dtab32[x] = nullptr; }
}

int plan_build_inner(lcpao_ctx *ctx);

// Builds the block plan once.  A geometry the block forms cannot hold (more
// than 32 orbitals on an atom, too many orbital rows or sphere images on a
// grid block, index ranges) gives LCPAO_ERANGE and marks the context
// (block_state 0), so that the callers switch to the pair-list kernels and
// the plan is not attempted again.
int plan_build(lcpao_ctx *ctx) {
    if (ctx->plan) return LCPAO_OK;
    if (ctx->block_state == 0) { ctx->err = ctx->block_why; return LCPAO_ERANGE; }
    if (ctx->maxno > 32) {
        ctx->err = "the block forms hold at most 32 orbitals per atom (this geometry has " + std::to_string(ctx->maxno) + ")";
    } else {
        const int rc = plan_build_inner(ctx);
        if (rc != LCPAO_ERANGE) {
            if (rc == LCPAO_OK) ctx->block_state = 1;
            return rc;
        }
    }
    ctx->block_state = 0;
    ctx->block_why = ctx->err;
    return LCPAO_ERANGE;
}

int plan_build_inner(lcpao_ctx *ctx) {
    const int N1 = ctx->ngrid[0], N2 = ctx->ngrid[1], N3 = ctx->ngrid[2];
    if (ctx->h_clist.empty()) { ctx->err = "split-precision forms need the cell lists"; return LCPAO_ESTATE; }
    auto *P = new BlockPlan();
    P->nb[0] = (N1 + BLK - 1) / BLK; P->nb[1] = (N2 + BLK - 1) / BLK; P->nb[2] = (N3 + BLK - 1) / BLK;
    P->nblocks = (int64_t)P->nb[0] * P->nb[1] * P->nb[2];
    const int64_t NB = P->nblocks;
    // bucket the grid-list entries by block: (atom, cell, point, column)
    struct Ent { int32_t atom, cell, k, t; };
    std::vector<int64_t> cnt(NB + 1, 0);
    auto blk_of = [&](int g, int &k) {
        const int n1 = g / (N2 * N3), n2 = (g / N3) % N2, n3 = g % N3;
        k = ((n1 % BLK) * BLK + (n2 % BLK)) * BLK + (n3 % BLK);
        return ((int64_t)(n1 / BLK) * P->nb[1] + (n2 / BLK)) * P->nb[2] + (n3 / BLK);
    };
    for (int64_t e = 0; e < ctx->total_gridn; e++) { int k; cnt[blk_of(ctx->h_glist[e], k) + 1]++; }
    for (int64_t b = 0; b < NB; b++) cnt[b + 1] += cnt[b];
    std::vector<Ent> ents(ctx->total_gridn);
    {
        std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
        for (int a = 0; a < ctx->natom; a++)
            for (int t = 0; t < ctx->gridn[a]; t++) {
                const int64_t e = ctx->gl_off[a] + t;
                int k;
                const int64_t b = blk_of(ctx->h_glist[e], k);
                ents[fill[b]++] = {a, ctx->h_clist[e], k, t};
            }
    }
    // spheres of each block, in (atom, cell) order
    std::vector<std::vector<std::pair<int32_t, int32_t>>> sph(NB);  // (atom, cell)
    std::vector<std::vector<int32_t>> rows0(NB);
    const int nthreads = (int)std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    {
        std::vector<std::thread> th;
        for (int w = 0; w < nthreads; w++)
            th.emplace_back([&, w] {
                for (int64_t b = w; b < NB; b += nthreads) {
                    std::sort(ents.begin() + cnt[b], ents.begin() + cnt[b + 1], [](const Ent &x, const Ent &y) {
                        return x.atom != y.atom ? x.atom < y.atom : (x.cell != y.cell ? x.cell < y.cell : x.k < y.k);
                    });
                    for (int64_t e = cnt[b]; e < cnt[b + 1]; e++)
                        if (sph[b].empty() || sph[b].back().first != ents[e].atom || sph[b].back().second != ents[e].cell)
                            sph[b].push_back({ents[e].atom, ents[e].cell});
                }
            });
        for (auto &x : th) x.join();
    }
    P->blocks.resize(NB);
    std::vector<int64_t> sph_first(NB + 1, 0);
    int64_t toff = 0, poff = 0, spoff = 0, roff = 0, soff = 0;
    for (int64_t b = 0; b < NB; b++) {
        int r = 0;
        rows0[b].resize(sph[b].size());
        for (size_t s = 0; s < sph[b].size(); s++) { rows0[b][s] = r; r += ctx->no[sph[b][s].first]; }
        const int R = std::max(ST, (r + ST - 1) / ST * ST);
        if (R > MAX_ROWS) {
            ctx->err = "a grid block is covered by more orbitals than the split-precision kernels hold (" + std::to_string(r) + ")";
            delete P;
            return LCPAO_ERANGE;
        }
        P->max_rows = std::max(P->max_rows, R);
        const int S = R / ST, nst = S * (S + 1) / 2;
        const int b1 = (int)(b / ((int64_t)P->nb[1] * P->nb[2])), b2 = (int)((b / P->nb[2]) % P->nb[1]), b3 = (int)(b % P->nb[2]);
        const int ns = (int)sph[b].size();
        if (roff + R > INT32_MAX) { ctx->err = "too many grid blocks for the split-precision plan"; delete P; return LCPAO_ERANGE; }
        if (poff + nst > INT32_MAX) { ctx->err = "too many super-tiles for the split-precision plan"; delete P; return LCPAO_ERANGE; }
        P->blocks[b] = {toff, spoff, R, b1 * BLK, b2 * BLK, b3 * BLK, ns, (int32_t)roff, (int32_t)soff, (int32_t)poff, r, 0};
        P->max_slots = std::max(P->max_slots, ns);
        if (ns * ns > MAX_SLOT_PAIRS) {
            ctx->err = "a grid block is covered by more sphere images than the split-precision kernels hold";
            delete P;
            return LCPAO_ERANGE;
        }
        soff += ns;
        toff += (int64_t)R * BLK_PTS;
        poff += nst;
        spoff += (int64_t)ns * ns;
        roff += R;
        for (int st0 = 0; st0 < nst; st0 += HB_WAVES * ST_PER_WAVE)
            P->items.push_back({(int32_t)b, st0, std::min(HB_WAVES * ST_PER_WAVE, nst - st0), 0});
        if (R <= RH_ROWS) {
            const int n8 = (nst + RH_WAVES - 1) / RH_WAVES;
            for (int k = 0, st0 = 0; k < n8; k++) {
                const int cnt = nst / n8 + (k < nst % n8 ? 1 : 0);
                P->items8.push_back({(int32_t)b, st0, cnt, 0});
                st0 += cnt;
            }
        } else {
            for (int st0 = 0; st0 < nst; st0 += HB_WAVES * ST_PER_WAVE)
                P->items_big.push_back({(int32_t)b, st0, std::min(HB_WAVES * ST_PER_WAVE, nst - st0), 0});
            P->max_rows_big = std::max(P->max_rows_big, R);
        }
        sph_first[b + 1] = sph_first[b] + (int64_t)sph[b].size();
    }
    P->tile_elems = toff;
    P->part_tiles = poff;
    P->nsph = sph_first[NB];
    // items by decreasing work (rows, then super-tiles) so that the largest
    // start first; the row classes are kept for the plan statistics
    std::stable_sort(P->items.begin(), P->items.end(), [&](const HItem &x, const HItem &y) {
        const int rx = P->blocks[x.block].rows, ry = P->blocks[y.block].rows;
        return rx != ry ? rx > ry : x.nst > y.nst;
    });
    for (size_t i = 0; i < P->items.size(); i++) {
        const int R = P->blocks[P->items[i].block].rows;
        if (P->classes.empty() || P->classes.back().rows != R) P->classes.push_back({R, (int32_t)i, 0});
        P->classes.back().count++;
    }
    deal_xcd(P, P->items, ctx->ngrid[0]);
    // a block's items are consecutive in its slab, so they run one after the
    // other on one XCD and its L2 serves the second its tiles
    deal_xcd(P, P->items8, ctx->ngrid[0]);
    // row descriptors: slot * 32 + orbital (-1 for padding rows)
    std::vector<int16_t> rowinfo(std::max<int64_t>(1, roff), -1);
    for (int64_t b = 0; b < NB; b++)
        for (size_t x = 0; x < sph[b].size(); x++)
            for (int i = 0; i < ctx->no[sph[b][x].first]; i++)
                rowinfo[P->blocks[b].row_off + rows0[b][x] + i] = (int16_t)(x * 32 + i);
    if (ctx->my_hsize > INT32_MAX) { ctx->err = "pair blocks exceed the split-precision index range"; delete P; return LCPAO_ERANGE; }
    std::vector<int16_t> slotno(std::max<int64_t>(1, soff), 0);
    for (int64_t b = 0; b < NB; b++)
        for (size_t x = 0; x < sph[b].size(); x++) slotno[P->blocks[b].slot_off + x] = (int16_t)ctx->no[sph[b][x].first];
    std::vector<SlotPair> spairs(std::max<int64_t>(1, spoff), SlotPair{-1, 0});
    std::vector<SlotPair> fpairs(std::max<int64_t>(1, spoff), SlotPair{-1, 0});  // blocks own disjoint ranges
    // sphere tables and point maps
    std::vector<int32_t> sb(P->nsph), sa(P->nsph), sr(P->nsph), smap((size_t)P->nsph * BLK_PTS, -1);
    std::vector<uint64_t> mask((size_t)P->nsph * (BLK_PTS / 64), 0);
    for (int64_t b = 0; b < NB; b++) {
        size_t s = 0;
        for (int64_t e = cnt[b]; e < cnt[b + 1]; e++) {
            while (sph[b][s].first != ents[e].atom || sph[b][s].second != ents[e].cell) s++;
            const int64_t gs = sph_first[b] + (int64_t)s;
            smap[(size_t)gs * BLK_PTS + ents[e].k] = ents[e].t;
            mask[(size_t)gs * 8 + ents[e].k / 64] |= 1ull << (ents[e].k % 64);
        }
        for (size_t q = 0; q < sph[b].size(); q++) {
            const int64_t gs = sph_first[b] + (int64_t)q;
            sb[gs] = (int32_t)b; sa[gs] = sph[b][q].first; sr[gs] = rows0[b][q];
        }
        // orbital rows present on each 4x4x4 sub-block (mask word = first
        // point index, bit = second * 8 + third within the block)
        int sub_rows[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (size_t q = 0; q < sph[b].size(); q++) {
            const uint64_t *m = &mask[(size_t)(sph_first[b] + (int64_t)q) * 8];
            for (int sc = 0; sc < 8; sc++) {
                const int s1 = sc >> 2, s2 = (sc >> 1) & 1, s3 = sc & 1;
                uint64_t sel = 0;
                for (int u = 0; u < 4; u++) sel |= 0xFull << ((4 * s2 + u) * 8 + 4 * s3);
                bool any = false;
                for (int u = 0; u < 4; u++) any = any || (m[4 * s1 + u] & sel) != 0;
                if (any) sub_rows[sc] += ctx->no[sph[b][q].first];
            }
        }
        for (int sc = 0; sc < 8; sc++) P->max_sub_rows = std::max(P->max_sub_rows, sub_rows[sc]);
    }
    // contributions of each computed pair
    const int C = ctx->cpy, W = 2 * C + 1;
    std::unordered_map<int64_t, int32_t> pair_of;
    pair_of.reserve(ctx->npair * 2);
    auto pkey = [&](int a, int b, int i, int j, int k) {
        return (((int64_t)a * ctx->natom + b) * W + (i + C)) * W * W + (int64_t)(j + C) * W + (k + C);
    };
    for (int64_t p = 0; p < ctx->npair; p++) {
        int ijk[3];
        cell_ijk3(C, ctx->pair_ncn[p], ijk);
        pair_of[pkey(ctx->pairs[p].a, ctx->pairs[p].b, ijk[0], ijk[1], ijk[2])] = (int32_t)p;
    }
    std::vector<int32_t> canon_pos(ctx->npair, -1);
    for (size_t c = 0; c < ctx->canon.size(); c++) canon_pos[ctx->canon[c]] = (int32_t)c;
    struct CRec { int32_t cpos; int32_t block, trans; int32_t x, y; };
    std::vector<std::vector<CRec>> recs(nthreads);
    std::vector<std::vector<int64_t>> olg_count(nthreads, std::vector<int64_t>(ctx->npair, 0));
    {
        std::vector<std::thread> th;
        for (int w = 0; w < nthreads; w++)
            th.emplace_back([&, w] {
                for (int64_t b = w; b < NB; b += nthreads) {
                    const int64_t s0 = sph_first[b], ns = sph_first[b + 1] - s0;
                    for (int64_t x = 0; x < ns; x++)
                        for (int64_t y = x; y < ns; y++) {
                            int ov = 0;
                            for (int m = 0; m < 8; m++) ov += __builtin_popcountll(mask[(s0 + x) * 8 + m] & mask[(s0 + y) * 8 + m]);
                            if (!ov) continue;
                            int ix[3], iy[3];
                            cell_ijk3(C, sph[b][x].second, ix);
                            cell_ijk3(C, sph[b][y].second, iy);
                            const int d0 = ix[0] - iy[0], d1 = ix[1] - iy[1], d2 = ix[2] - iy[2];
                            if (std::abs(d0) > C || std::abs(d1) > C || std::abs(d2) > C) continue;
                            const int ax = sph[b][x].first, ay = sph[b][y].first;
                            auto fx = pair_of.find(pkey(ax, ay, d0, d1, d2));
                            auto fy = pair_of.find(pkey(ay, ax, -d0, -d1, -d2));
                            const int32_t pxy = fx == pair_of.end() ? -1 : fx->second;
                            const int32_t pyx = fy == pair_of.end() ? -1 : fy->second;
                            if (pxy >= 0) olg_count[w][pxy] += ov;
                            if (pyx >= 0 && x != y) olg_count[w][pyx] += ov;
                            const int64_t sp0 = P->blocks[b].sp_off;
                            if (pxy >= 0)
                                fpairs[sp0 + x * ns + y] = SlotPair{(int32_t)ctx->pairs[pxy].hoff, ctx->pairs[pxy].nob};
                            if (pyx >= 0)
                                fpairs[sp0 + y * ns + x] = SlotPair{(int32_t)ctx->pairs[pyx].hoff, ctx->pairs[pyx].nob};
                            int32_t pc = -1, trans = 0;
                            if (pxy >= 0 && canon_pos[pxy] >= 0) pc = pxy;
                            else if (pyx >= 0 && canon_pos[pyx] >= 0) { pc = pyx; trans = 1; }
                            if (pc < 0) continue;
                            recs[w].push_back({canon_pos[pc], (int32_t)b, trans, (int32_t)x, (int32_t)y});
                        }
                }
            });
        for (auto &x : th) x.join();
    }
    for (int64_t p = 0; p < ctx->npair; p++) {
        int64_t n = 0;
        for (int w = 0; w < nthreads; w++) n += olg_count[w][p];
        if (n != ctx->numolg[p]) {
            ctx->err = "overlap lists are inconsistent with the grid lists (pair " + std::to_string(p) + ")";
            delete P;
            return LCPAO_EINVAL;
        }
    }
    std::vector<CRec> all;
    for (auto &v : recs) all.insert(all.end(), v.begin(), v.end());
    // the reduction's launch order: pairs with the most contributions first
    // (on-site pairs have several times the average), so that the longest
    // workgroups do not start last; ties in canonical order
    const int64_t ncan = (int64_t)ctx->canon.size();
    std::vector<int32_t> ncontrib(ncan, 0), lorder(ncan), lpos(ncan), red_canon(ncan);
    for (const CRec &r : all) ncontrib[r.cpos]++;
    for (int64_t c = 0; c < ncan; c++) lorder[c] = (int32_t)c;
    std::stable_sort(lorder.begin(), lorder.end(), [&](int32_t x, int32_t y) { return ncontrib[x] > ncontrib[y]; });
    P->red_count.resize(ncan);
    for (int64_t n = 0; n < ncan; n++) {
        lpos[lorder[n]] = (int32_t)n;
        red_canon[n] = ctx->canon[lorder[n]];
        P->red_count[n] = ncontrib[lorder[n]];
    }
    // within a pair: by the reduction's class (trans, first row mod 4), then by block
    auto rclass = [&](const CRec &r) { return r.trans * 4 + (rows0[r.block][r.x] & 3); };
    std::stable_sort(all.begin(), all.end(), [&](const CRec &x, const CRec &y) {
        if (x.cpos != y.cpos) return lpos[x.cpos] < lpos[y.cpos];
        const int cx = rclass(x), cy = rclass(y);
        return cx != cy ? cx < cy : x.block < y.block;
    });
    std::vector<Contrib> contrib(all.size());
    std::vector<int32_t> cpair(all.size());
    std::vector<int32_t> coff(ctx->canon.size() + 1, 0);
    // contribution records, in launch order: where the pair's sub-block lies
    // in the block's super-tile partials
    for (size_t i = 0; i < all.size(); i++) {
        const int64_t b = all[i].block;
        const BlockDev &B = P->blocks[b];
        cpair[i] = ctx->canon[all[i].cpos];
        contrib[i] = Contrib{(int64_t)B.st_off * ST * ST, (int16_t)rows0[b][all[i].x], (int16_t)rows0[b][all[i].y],
                             (int16_t)(B.rows / ST), (int16_t)all[i].trans};
        coff[lpos[all[i].cpos] + 1]++;
        const PairDev &Q = ctx->pairs[cpair[i]];
        const int64_t sp = B.sp_off + (int64_t)all[i].x * B.nslot + all[i].y;
        spairs[sp] = SlotPair{(int32_t)Q.hoff, Q.nob | (all[i].trans << 6)};
    }
    P->part_elems = P->part_tiles * ST * ST;
    P->ncontrib = (int64_t)all.size();
    for (int32_t p : ctx->canon)
        P->red_threads = std::max(P->red_threads, hblk_reduce_width(ctx->pairs[p].noa, ctx->pairs[p].nob));
    for (size_t c = 0; c < ctx->canon.size(); c++) coff[c + 1] += coff[c];
    // Force3 items: the S x S super-tiles of a block in units of two of one
    // row ((si, 2 u), (si, 2 u + 1); U = ceil(S / 2) units per row), FB_WAVES
    // units per item
    for (int64_t b = 0; b < NB; b++) {
        const int S = P->blocks[b].rows / ST, U = (S + 1) / 2;
        for (int u0 = 0; u0 < S * U; u0 += FB_WAVES)
            P->fitems.push_back({(int32_t)b, u0, std::min(FB_WAVES, S * U - u0), 0});
    }
    deal_xcd(P, P->fitems, ctx->ngrid[0]);
    std::vector<int32_t> fbase(std::max<size_t>(1, P->fitems.size()));
    std::vector<std::vector<int32_t>> fatom(ctx->natom);
    for (size_t i = 0; i < P->fitems.size(); i++) {
        const int64_t b = P->fitems[i].block;
        fbase[i] = (int32_t)P->fout;
        for (size_t x = 0; x < sph[b].size(); x++) fatom[sph[b][x].first].push_back((int32_t)(P->fout + (int64_t)x));
        P->fout += (int64_t)sph[b].size();
    }
    if (P->fout > INT32_MAX) { ctx->err = "Force3 sums exceed the split-precision index range"; delete P; return LCPAO_ERANGE; }
    std::vector<int32_t> fatom_off(ctx->natom + 1, 0), fatom_idx;
    fatom_idx.reserve(std::max<int64_t>(1, P->fout));
    for (int a = 0; a < ctx->natom; a++) {
        fatom_idx.insert(fatom_idx.end(), fatom[a].begin(), fatom[a].end());
        fatom_off[a + 1] = (int32_t)fatom_idx.size();
    }
    int rc;
    ctx->plan = P;
    if ((rc = up(ctx, &P->d_blocks, P->blocks)) || (rc = up(ctx, &P->d_items, P->items)) ||
        (rc = up(ctx, &P->d_items8, P->items8)) || (rc = up(ctx, &P->d_items_big, P->items_big)) ||
        (rc = up(ctx, &P->d_sph_block, sb)) || (rc = up(ctx, &P->d_sph_atom, sa)) || (rc = up(ctx, &P->d_sph_row0, sr)) ||
        (rc = up(ctx, &P->d_sph_map, smap)) || (rc = up(ctx, &P->d_contrib, contrib)) || (rc = up(ctx, &P->d_contrib_off, coff)) ||
        (rc = up(ctx, &P->d_red_canon, red_canon)) ||
        (rc = up(ctx, &P->d_contrib_pair, cpair)) || (rc = up(ctx, &P->d_rowinfo, rowinfo)) ||
        (rc = up(ctx, &P->d_spair, spairs)) ||
        (rc = up(ctx, &P->d_slotno, slotno)) || (rc = up(ctx, &P->d_fpair, fpairs)) ||
        (rc = up(ctx, &P->d_fitems, P->fitems)) ||
        (rc = up(ctx, &P->d_fbase, fbase)) || (rc = up(ctx, &P->d_fatom_off, fatom_off)) ||
        (rc = up(ctx, &P->d_fatom_idx, fatom_idx))) {
        ctx->plan = nullptr;
        plan_free(P);
        return rc;
    }
    return LCPAO_OK;
}

int split_tables(lcpao_ctx *ctx, int precision) {
    int rc = plan_build(ctx);
    if (rc) return rc;
    BlockPlan *P = ctx->plan;
    const bool bf = precision == LCPAO_BF16X3;
    if (bf ? P->bf16_ready : P->f32_ready) return LCPAO_OK;
    if (!ctx->d_deff32) SPLIT_TRY(ctx, hipMalloc(&ctx->d_deff32, std::max<int64_t>(1, ctx->my_hsize) * sizeof(float)));
    if (bf) {
        if (!P->d_hi) {
            const size_t bytes = std::max<int64_t>(1, P->tile_elems) * sizeof(uint16_t);
            SPLIT_TRY(ctx, hipMalloc(&P->d_hi, bytes));
            SPLIT_TRY(ctx, hipMalloc(&P->d_lo, bytes));
            SPLIT_TRY(ctx, hipMalloc(&P->d_thi, bytes));
            SPLIT_TRY(ctx, hipMalloc(&P->d_tlo, bytes));
        }
        SPLIT_TRY(ctx, hipMemsetAsync(P->d_hi, 0, P->tile_elems * sizeof(uint16_t), ctx->stream));
        SPLIT_TRY(ctx, hipMemsetAsync(P->d_lo, 0, P->tile_elems * sizeof(uint16_t), ctx->stream));
        SPLIT_TRY(ctx, hipMemsetAsync(P->d_thi, 0, P->tile_elems * sizeof(uint16_t), ctx->stream));
        SPLIT_TRY(ctx, hipMemsetAsync(P->d_tlo, 0, P->tile_elems * sizeof(uint16_t), ctx->stream));
    } else {
        if (!P->d_t32) SPLIT_TRY(ctx, hipMalloc(&P->d_t32, std::max<int64_t>(1, P->tile_elems) * sizeof(float)));
        SPLIT_TRY(ctx, hipMemsetAsync(P->d_t32, 0, P->tile_elems * sizeof(float), ctx->stream));
    }
    const int64_t nthr = P->nsph * BLK_PTS;
    if (nthr > 0)
        hipLaunchKernelGGL(k_tiles, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, ctx->stream, P->nsph, P->d_sph_block,
                           P->d_sph_atom, P->d_sph_row0, P->d_sph_map, P->d_blocks, ctx->d_tab_off, ctx->d_no, ctx->d_tab,
                           precision, P->d_hi, P->d_lo, P->d_t32, P->d_thi, P->d_tlo);
    SPLIT_TRY(ctx, hipGetLastError());
    if (!P->d_part) SPLIT_TRY(ctx, hipMalloc(&P->d_part, std::max<int64_t>(1, P->part_elems) * sizeof(float)));
    SPLIT_TRY(ctx, hipStreamSynchronize(ctx->stream));
    (bf ? P->bf16_ready : P->f32_ready) = true;
    return LCPAO_OK;
}

// Autotuner of the block kernels' candidates `launch(v)` for which fits(v)
// holds: each once untimed, then TUNE_ROUNDS timed launches per candidate in
// alternating order, keeping each one's fastest (a single launch each is
// within the noise of a GPU); the fastest candidate is kept in
// ctx->variant[op], and a candidate that does not fit reads -1 ms.
constexpr int TUNE_ROUNDS = 4;
template <class F, class FITS>
hipError_t tune_block_kernel(lcpao_ctx *ctx, int op, F launch, FITS fits) {
    const int nc = TUNE_NCAND[op];
    for (int v = 0; v < nc; v++) ctx->variant_ms[op][v] = -1.f;
    hipEvent_t e0, e1;
    hipError_t e = hipEventCreate(&e0);
    if (e != hipSuccess) return e;
    if ((e = hipEventCreate(&e1)) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return e;
    }
    for (int v = 0; v < nc && e == hipSuccess; v++)
        if (fits(v)) e = launch(v);
    for (int r = 0; r < TUNE_ROUNDS && e == hipSuccess; r++)
        for (int x = 0; x < nc && e == hipSuccess; x++) {
            const int v = r % 2 ? nc - 1 - x : x;
            if (!fits(v)) continue;
            float ms = 0.f;
            if ((e = hipEventRecord(e0, ctx->stream)) != hipSuccess || (e = launch(v)) != hipSuccess ||
                (e = hipEventRecord(e1, ctx->stream)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess ||
                (e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess)
                break;
            float &m = ctx->variant_ms[op][v];
            if (r == 0 || ms < m) m = ms;
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return e;
    int best = 0;
    for (int v = 1; v < nc; v++)
        if (fits(v) && ctx->variant_ms[op][v] < ctx->variant_ms[op][best]) best = v;
    ctx->variant[op] = best;
    return hipSuccess;
}

int split_hamiltonian(lcpao_ctx *ctx, int nch, const double *const *d_V, double *const *d_H, int precision) {
    int rc = split_tables(ctx, precision);
    if (rc) return rc;
    BlockPlan *P = ctx->plan;
    const bool bf = precision == LCPAO_BF16X3;
    for (int c = 0; c < nch; c++) {
        if (!d_V[c] || !d_H[c]) { ctx->err = "null channel pointer"; return LCPAO_EINVAL; }
        if (!P->items.empty()) {
            KTimer kt(ctx, KID_HAM);
            // one launch for all items: the kernels run one workgroup per CU
            // (register-bound) whatever R is, so LDS and prefetch depth are
            // sized for the largest block at no cost in occupancy
            const unsigned ni = (unsigned)P->items.size();
            if (bf) {
                const BlockArgs one = block_args(ctx, d_V[c], nullptr);
                SPLIT_TRY(ctx, launch_block_kernel(false, ctx->stream, P->max_rows, 0, P->d_items, (int32_t)ni, nullptr,
                                                   one, nullptr, ni));
            } else {
                SPLIT_TRY(ctx, launch_f32_block_kernel(false, ctx->stream, P->max_rows, P->max_slots, P->d_items,
                                                       (int32_t)ni, nullptr, block_args(ctx, d_V[c], nullptr)));
            }
        }
        if (ctx->ncanon > 0) {
            KTimer kt(ctx, KID_HRED);
            hipLaunchKernelGGL(k_hblk_reduce, dim3((unsigned)ctx->ncanon), dim3(P->red_threads), 0, ctx->stream, P->d_red_canon,
                               ctx->ncanon, P->d_contrib_off, P->d_contrib, ctx->d_pairs, P->d_part, ctx->gridvol,
                               d_H[c]);
        }
        SPLIT_TRY(ctx, hipGetLastError());
    }
    SPLIT_TRY(ctx, hipStreamSynchronize(ctx->stream));
    collect_timings(ctx);
    return LCPAO_OK;
}

int fp32_tables(lcpao_ctx *ctx, bool deriv) {
    const int64_t total = ctx->tab_off[ctx->natom];
    const unsigned nb = (unsigned)((total + 255) / 256);
    if (!ctx->d_tab32) {
        SPLIT_TRY(ctx, hipMalloc(&ctx->d_tab32, std::max<int64_t>(1, total) * sizeof(float)));
        if (total) hipLaunchKernelGGL(k_to_f32, dim3(nb), dim3(256), 0, ctx->stream, total, ctx->d_tab, ctx->d_tab32);
    }
    if (deriv)
        for (int x = 0; x < 3; x++)
            if (!ctx->d_dtab32[x]) {
                SPLIT_TRY(ctx, hipMalloc(&ctx->d_dtab32[x], std::max<int64_t>(1, total) * sizeof(float)));
                if (total) hipLaunchKernelGGL(k_to_f32, dim3(nb), dim3(256), 0, ctx->stream, total, ctx->d_dtab[x], ctx->d_dtab32[x]);
            }
    if (!ctx->d_deff32) SPLIT_TRY(ctx, hipMalloc(&ctx->d_deff32, std::max<int64_t>(1, ctx->my_hsize) * sizeof(float)));
    SPLIT_TRY(ctx, hipGetLastError());
    SPLIT_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return LCPAO_OK;
}


// ----------------------------------------------------------------------------
// block density: rho_k = sum_{si<=sj} Phi_si^T W[si][sj] Phi_sj  (bf16x3)
//   W = D + D^T on off-diagonal super-tiles, (D + D^T)/2 on diagonal ones,
//   gathered per item from the folded pair blocks through the block's
//   slot-pair table (no W plane in HBM).
// ----------------------------------------------------------------------------
__global__ void k_fill_f64(int64_t n, double *__restrict__ x, double v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

__global__ void k_fill_i32(int64_t n, int32_t *__restrict__ x, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

// host wrappers for the batched form (fills instead of memsets: they may be
// captured in a graph)
void launch_fill_f64(hipStream_t stream, int64_t n, double *x, double v) {
    if (n > 0) hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, n, x, v);
}
void launch_fill_i32(hipStream_t stream, int64_t n, int32_t *x, int32_t v) {
    if (n > 0) hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, n, x, v);
}
void launch_hblk_reduce_batch(hipStream_t stream, int nstruct, const BatchRed *d_red, const int32_t *d_off,
                              int total_canon, int threads, const int32_t *d_perm) {
    if (nstruct > 0 && total_canon > 0)
        hipLaunchKernelGGL(k_hblk_reduce_batch, dim3((unsigned)total_canon), dim3(threads), 0, stream, d_red, d_off,
                           nstruct, d_perm);
}
void launch_fill_batch(hipStream_t stream, int nstruct, const BatchFill *d_fill, int64_t max_n) {
    if (nstruct > 0 && max_n > 0)
        hipLaunchKernelGGL(k_fill_batch, dim3((unsigned)((max_n + 255) / 256), (unsigned)nstruct), dim3(256), 0, stream,
                           d_fill);
}
void launch_hblk_reduce(const lcpao_ctx *ctx, hipStream_t stream, double *H) {
    const BlockPlan *P = ctx->plan;
    if (ctx->ncanon > 0)
        hipLaunchKernelGGL(k_hblk_reduce, dim3((unsigned)ctx->ncanon), dim3(P->red_threads), 0, stream, P->d_red_canon, ctx->ncanon,
                           P->d_contrib_off, P->d_contrib, ctx->d_pairs, P->d_part, ctx->gridvol, H);
}

constexpr int RB_BYTES_PER_ROW = 2 * CHUNK * 2;   // LDS bytes per row: hi + lo point fragments
constexpr int RB_WAVES = 16;                       // waves of the density workgroup, one super-tile each
constexpr int RB_SUMS_OFFSET = MAX_ROWS * RB_BYTES_PER_ROW;  // per-wave point sums, after the chunk planes
constexpr int RB_AREA_BYTES = RB_SUMS_OFFSET + RB_WAVES * BLK_PTS * 4;
constexpr int RB_IMAGE_BYTES = HB_WAVES * 8 * 2 * 512 * 2;    // W images of the 8-wave (batched) form
static_assert(RB_AREA_BYTES <= RB_IMAGE_BYTES, "the 8-wave form's LDS covers the planes and sums");
static_assert(RB_WAVES * (16 / RB_WAVES) == HB_WAVES * 2, "an item's super-tiles fill the density waves");
// Double-buffered form of the 16-wave kernel: two chunk-plane buffers (chunk
// c is staged into buffer c % 2 while the other holds chunk c - 1, so one
// barrier per chunk suffices), then the per-chunk wave sums (two buffers of
// [wave][32 points]) and the block's 512 point sums.
constexpr int RB_DB_BUF_BYTES = MAX_ROWS * RB_BYTES_PER_ROW;
constexpr int RB_DB_WSUM_OFFSET = 2 * RB_DB_BUF_BYTES;
constexpr int RB_DB_PSUM_OFFSET = RB_DB_WSUM_OFFSET + 2 * RB_WAVES * CHUNK * 4;
constexpr int RB_DB_AREA_BYTES = RB_DB_PSUM_OFFSET + BLK_PTS * 4;


// W fragments of one super-tile (si, sj) of a block for the 16-wave density
// kernels, gathered straight into the A-operand layout: fragment (rt, ks) of
// lane l holds W[16 rt + l % 16][32 ks + 8 (l / 16) + e], e < 8, split into
// bf16 hi and lo.  W element = mask * Deff[one index] (on-site blocks hold
// D + D^T, so every element is a single load), halved on a diagonal
// super-tile (whose mirrored half is not computed), zero for padding rows,
// absent pairs or a wave without a super-tile (keep = false); 32 loads in
// flight per batch, the batches kept apart so that the compiler does not
// hoist all 64 loads.
__device__ __forceinline__ void gather_w16(const GLOBAL_AS float *__restrict__ Deff, const int16_t *s_rinfo,
                                           const SlotPair *s_sp, int nslot, int si, int sj, bool keep, int l16, int q,
                                           bf16x8 (&wh)[4][2], bf16x8 (&wl)[4][2]) {
    constexpr int GB = 4;
    const float wdiag = si == sj ? 0.5f : 1.f;
    const float kf = keep ? 1.f : 0.f;
#pragma unroll
    for (int b2 = 0; b2 < 8; b2 += GB) {
        float dv[GB][8];
        uint32_t okbits = 0;
#pragma unroll
        for (int r2 = 0; r2 < GB; r2++) {
            const int rt = (b2 + r2) >> 1, ks = (b2 + r2) & 1;
            const int rinf = s_rinfo[si * ST + 16 * rt + l16];
            const int ra = rinf >> 5, ri = rinf & 31;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int cinf = s_rinfo[sj * ST + 32 * ks + 8 * q + e];
                const int cb = cinf >> 5, cj = cinf & 31;
                const bool valid = rinf >= 0 && cinf >= 0;
                const SlotPair P = s_sp[valid ? min(ra, cb) * nslot + max(ra, cb) : 0];
                const bool sw = ra > cb;  // lower triangle of a diagonal super-tile: W is symmetric
                const int i = sw ? cj : ri, j = sw ? ri : cj;
                const bool ok = valid && P.hoff >= 0;
                const int ld = P.meta & 63;
                const int x1 = (P.meta >> 6) ? j * ld + i : i * ld + j;
                dv[r2][e] = Deff[ok ? P.hoff + x1 : 0];  // unconditional: loads issue back to back
                okbits |= (ok ? 1u : 0u) << (r2 * 8 + e);
            }
        }
#pragma unroll
        for (int r2 = 0; r2 < GB; r2++) {
            const int rt = (b2 + r2) >> 1, ks = (b2 + r2) & 1;
            u16x8 wh8, wl8;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                // D is finite, so masking by a product keeps the load unconditional
                const float w = dv[r2][e] * (((okbits >> (r2 * 8 + e)) & 1u) ? wdiag * kf : 0.f);
                const __bf16 hb = (__bf16)w;
                wh8[e] = __builtin_bit_cast(uint16_t, hb);
                wl8[e] = __builtin_bit_cast(uint16_t, (__bf16)(w - (float)hb));
            }
            wh[rt][ks] = __builtin_bit_cast(bf16x8, wh8);
            wl[rt][ks] = __builtin_bit_cast(bf16x8, wl8);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Density of one item: W waves, each owning PER = 16 / W of the item's
// super-tiles.  With W = 16 a wave holds the W fragments of one super-tile
// (64 VGPRs), so four waves fit per SIMD.  DB (W = 16 only): double-buffered
// chunk planes with one barrier per chunk; the 16 wave sums of a chunk's
// points are added (in wave order, as in the single-buffered form) by one
// wave after the next barrier, so both forms give the same bits.
template <int NF, int W, bool DB>
__device__ __forceinline__ void rblk_item(const HItem it, const BlockArgs &A, unsigned char *lds_raw) {
    constexpr int PER = 16 / W;
    static_assert(!DB || W == RB_WAVES, "the double-buffered form has one super-tile per wave");
    const BlockDev B = as_global(A.blocks)[it.block];
    const GLOBAL_AS uint16_t *__restrict__ thi = as_global(A.thi);
    const GLOBAL_AS uint16_t *__restrict__ tlo = as_global(A.tlo);
    const GLOBAL_AS int16_t *__restrict__ rowinfo = as_global(A.rowinfo);
    const GLOBAL_AS SlotPair *__restrict__ spair = as_global(A.spair);
    const GLOBAL_AS float *__restrict__ Deff = as_global(A.Deff);
    double *__restrict__ rho = A.rho;
    const int n1 = A.n1, n2 = A.n2, n3 = A.n3;
    const int R = B.rows, S = R / ST, G = R / 16, G32 = R / 32;
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63, l16 = lane & 15, q = lane >> 4;
    u16x8 *TH = reinterpret_cast<u16x8 *>(lds_raw);
    u16x8 *TL = TH + G * 64;
    // LDS: the chunk planes, the per-wave point sums, then the block's row
    // descriptors and slot pairs
    float *rcw = reinterpret_cast<float *>(lds_raw + RB_SUMS_OFFSET);  // [wave][512 points]
    float *wsum = reinterpret_cast<float *>(lds_raw + RB_DB_WSUM_OFFSET);  // DB: [chunk % 2][wave][32 points]
    float *psum = reinterpret_cast<float *>(lds_raw + RB_DB_PSUM_OFFSET);  // DB: [512 points]
    int16_t *s_rinfo =
        reinterpret_cast<int16_t *>(lds_raw + (DB ? RB_DB_AREA_BYTES : PER == 1 ? RB_AREA_BYTES : RB_IMAGE_BYTES));
    SlotPair *s_sp = reinterpret_cast<SlotPair *>(s_rinfo + MAX_ROWS);
    for (int x = tid; x < R; x += 64 * W) s_rinfo[x] = rowinfo[B.row_off + x];
    for (int x = tid; x < B.nslot * B.nslot; x += 64 * W) s_sp[x] = spair[B.sp_off + x];
    // W fragments of this wave's super-tiles (A operand: W[16 rt + l16][32 ks + 8 q + e])
    bf16x8 wh[PER][4][2], wl[PER][4][2];
    int si[PER], sj[PER];
    bool has[PER];
    // the first two chunks are fetched into L2 while W is gathered
    TouchN<64 * W> t0 = touch_chunk_n<64 * W>(thi, tlo, B.tile_off, R, tid);
    TouchN<64 * W> t1 = touch_chunk_n<64 * W>(thi, tlo, B.tile_off + (int64_t)G * FRAG, R, tid);
    __syncthreads();
#pragma unroll
    for (int h = 0; h < PER; h++) {
        has[h] = wave + h * W < it.nst;
        si[h] = sj[h] = 0;
        if (has[h]) st_decode(it.st0 + wave + h * W, S, si[h], sj[h]);
    }
    if constexpr (PER == 1) {
        gather_w16(Deff, s_rinfo, s_sp, B.nslot, si[0], sj[0], has[0], l16, q, wh[0], wl[0]);
    } else {
        // two super-tiles per wave: W of each is gathered into a per-wave LDS
        // image laid out as the A-operand fragments [rt * 2 + ks][hi, lo][lane][8]
        // and read back into registers, one super-tile at a time (the image of
        // the second one is built while the first one's fragments are live).
        // lane = (column group cg = lane / 8, row rr = lane % 8): the lane
        // fills W[8 t8 + rr][8 cg .. 8 cg + 7]; 32 loads per batch.
        static_assert(W * 8 * 2 * FRAG * 2 <= RB_IMAGE_BYTES, "W images exceed their LDS area");
        uint16_t *img = reinterpret_cast<uint16_t *>(lds_raw) + wave * (8 * 2 * FRAG);
        const int rr = lane & 7, cg = lane >> 3;
#pragma unroll
        for (int h = 0; h < PER; h++) {
            if (has[h]) {
                const float wdiag = si[h] == sj[h] ? 0.5f : 1.f;
                int ci[8];
#pragma unroll
                for (int e = 0; e < 8; e++) ci[e] = s_rinfo[sj[h] * ST + cg * 8 + e];
#pragma unroll 1
                for (int tb = 0; tb < 8; tb += 4) {
                    float dv[4][8];
                    uint32_t okbits = 0;
#pragma unroll
                    for (int r4 = 0; r4 < 4; r4++) {
                        const int rinf = s_rinfo[si[h] * ST + (tb + r4) * 8 + rr];
                        const int ra = rinf >> 5, ri = rinf & 31;
#pragma unroll
                        for (int e = 0; e < 8; e++) {
                            const int cb = ci[e] >> 5, cj = ci[e] & 31;
                            const bool valid = rinf >= 0 && ci[e] >= 0;
                            const SlotPair P = s_sp[valid ? min(ra, cb) * B.nslot + max(ra, cb) : 0];
                            const bool sw = ra > cb;
                            const int i = sw ? cj : ri, j = sw ? ri : cj;
                            const bool ok = valid && P.hoff >= 0;
                            const int ld = P.meta & 63;
                            const int x1 = (P.meta >> 6) ? j * ld + i : i * ld + j;
                            dv[r4][e] = Deff[ok ? P.hoff + x1 : 0];
                            okbits |= (ok ? 1u : 0u) << (r4 * 8 + e);
                        }
                    }
#pragma unroll
                    for (int r4 = 0; r4 < 4; r4++) {
                        u16x8 wh8, wl8;
#pragma unroll
                        for (int e = 0; e < 8; e++) {
                            const float w = dv[r4][e] * (((okbits >> (r4 * 8 + e)) & 1u) ? wdiag : 0.f);
                            const __bf16 hb = (__bf16)w;
                            wh8[e] = __builtin_bit_cast(uint16_t, hb);
                            wl8[e] = __builtin_bit_cast(uint16_t, (__bf16)(w - (float)hb));
                        }
                        const int row = (tb + r4) * 8 + rr;
                        const int frag = (row / 16) * 2 + cg / 4, at = ((cg % 4) * 16 + row % 16) * 8;
                        *reinterpret_cast<u16x8 *>(&img[(frag * 2 + 0) * FRAG + at]) = wh8;
                        *reinterpret_cast<u16x8 *>(&img[(frag * 2 + 1) * FRAG + at]) = wl8;
                    }
                }
            }
            __syncthreads();
            if (has[h]) {
                const u16x8 *im = reinterpret_cast<const u16x8 *>(img);
#pragma unroll
                for (int rt = 0; rt < 4; rt++)
#pragma unroll
                    for (int ks = 0; ks < 2; ks++) {
                        wh[h][rt][ks] = __builtin_bit_cast(bf16x8, im[((rt * 2 + ks) * 2 + 0) * 64 + lane]);
                        wl[h][rt][ks] = __builtin_bit_cast(bf16x8, im[((rt * 2 + ks) * 2 + 1) * 64 + lane]);
                    }
            }
            __syncthreads();
        }
    }
    u16x8 fh[NF], fl[NF];
    // fragment f of this wave; past the last one a wave reloads (and restages)
    // its first fragment, which its L1 already holds, rather than all such
    // waves fetching the last fragment together
    auto frag_of = [&](int f) {
        const int g = wave + f * W;
        return g < G ? g : min(wave, G - 1);
    };
    auto prefetch = [&](int c) {
        const GLOBAL_AS uint16_t *sh = thi + B.tile_off + (int64_t)c * G * FRAG;
        const GLOBAL_AS uint16_t *sl = tlo + B.tile_off + (int64_t)c * G * FRAG;
#pragma unroll
        for (int f = 0; f < NF; f++) {  // unconditional loads: see touch_chunk
            const int g = frag_of(f);
            fh[f] = *reinterpret_cast<const GLOBAL_AS u16x8 *>(sh + (g * 64 + lane) * 8);
            fl[f] = *reinterpret_cast<const GLOBAL_AS u16x8 *>(sl + (g * 64 + lane) * 8);
        }
    };
    // DB: chunk c + 1 is written into the other buffer while chunk c is read;
    // the barrier of step c also orders the writes of step c + 1 after every
    // wave's reads of chunk c - 1 (they precede that barrier), so the
    // barrier at the top of the single-buffered step is not needed
    auto add_wave_sums = [&](int c) {  // DB: the 16 wave sums of chunk c's points
        if (wave == c % W && lane < CHUNK) {
            const float *src = wsum + (c & 1) * W * CHUNK;
            float sum = 0.f;
            // in wave order, four loads in flight (the rest of the kernel
            // holds 120 VGPRs: sixteen at once would spill)
#pragma unroll
            for (int w0 = 0; w0 < W; w0 += 4) {
                float v[4];
#pragma unroll
                for (int k = 0; k < 4; k++) v[k] = src[(w0 + k) * CHUNK + lane];
#pragma unroll
                for (int k = 0; k < 4; k++) sum += v[k];
                __builtin_amdgcn_sched_barrier(0);
            }
            psum[c * CHUNK + lane] = sum;
        }
    };
    auto step = [&](int c, TouchN<64 * W> &slot) {
        // the buffer offset (72 KB) exceeds the LDS instructions' immediate
        // offsets: kept opaque in an SGPR, it costs one add per access instead
        // of a second set of address registers
        int bsel = DB ? (c & 1) : 0;
        if constexpr (DB) asm volatile("" : "+s"(bsel));
        u16x8 *BH = TH + bsel * (RB_DB_BUF_BYTES / 16);
        u16x8 *BL = BH + G * 64;
        if constexpr (!DB) __syncthreads();
#pragma unroll
        for (int f = 0; f < NF; f++) {  // branch-free, as in k_hblk_bf
            const int g = frag_of(f);
            BH[g * 64 + lane] = fh[f];
            BL[g * 64 + lane] = fl[f];
        }
        release_n(slot);
        __syncthreads();
        if constexpr (DB) {
            if (c > 0) add_wave_sums(c - 1);
        }
        prefetch(min(c + 1, NCHUNK - 1));
        slot = touch_chunk_n<64 * W>(thi, tlo, B.tile_off + (int64_t)min(c + 2, NCHUNK - 1) * G * FRAG, R, tid);
        // per 16-point group: a lane's partial sum over its super-tiles; the
        // four lanes of a point then add into its LDS sum (no shuffles)
#pragma unroll 1
        for (int pt = 0; pt < 2; pt++) {
            float part = 0.f;
#pragma unroll
            for (int h = 0; h < PER; h++) {
                if (!has[h]) continue;
                bf16x8 bh[2], bl[2];
#pragma unroll
                for (int ks = 0; ks < 2; ks++) {
                    const int fi = pt * G32 + sj[h] * 2 + ks;
                    bh[ks] = __builtin_bit_cast(bf16x8, BH[fi * 64 + lane]);
                    bl[ks] = __builtin_bit_cast(bf16x8, BL[fi * 64 + lane]);
                }
#pragma unroll
                for (int rt = 0; rt < 4; rt++) {
                    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int ks = 0; ks < 2; ks++) {
                        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wl[h][rt][ks], bh[ks], acc, 0, 0, 0);
                        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh[h][rt][ks], bl[ks], acc, 0, 0, 0);
                        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh[h][rt][ks], bh[ks], acc, 0, 0, 0);
                    }
                    // acc[v] = Q[row si*64 + 16 rt + 4 q + v][point 16 pt + l16]; Phi at the
                    // same four rows and point: 8 contiguous bytes of a point fragment
                    const int fi = pt * G32 + si[h] * 2 + (rt >> 1);
                    const int off = (fi * 64 + (2 * (rt & 1) + (q >> 1)) * 16 + l16) * 8 + 4 * (q & 1);
                    const u16x4 ph = *reinterpret_cast<const u16x4 *>(reinterpret_cast<const uint16_t *>(BH) + off);
                    const u16x4 pl = *reinterpret_cast<const u16x4 *>(reinterpret_cast<const uint16_t *>(BL) + off);
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        part = fmaf(bf_bits_to_f32(ph[v]) + bf_bits_to_f32(pl[v]), acc[v], part);
                }
            }
            // sum over the four lanes of each point (lanes 16 apart) with the
            // gfx950 lane swaps: after both, every lane holds its point's sum
            const auto s32 = __builtin_amdgcn_permlane32_swap(__float_as_uint(part), __float_as_uint(part), false, false);
            part = __uint_as_float(s32[0]) + __uint_as_float(s32[1]);
            const auto s16 = __builtin_amdgcn_permlane16_swap(__float_as_uint(part), __float_as_uint(part), false, false);
            part = __uint_as_float(s16[0]) + __uint_as_float(s16[1]);
            if (q == 0) {
                if constexpr (DB) wsum[(c & 1) * W * CHUNK + wave * CHUNK + pt * 16 + l16] = has[0] ? part : 0.f;
                else rcw[wave * BLK_PTS + c * CHUNK + pt * 16 + l16] = has[0] ? part : 0.f;
            }
        }
    };
    prefetch(0);
    // fully unrolled: with no loop back-edge the compiler's waits count the
    // loads exactly and never wait for a touch issued in the step before
#pragma unroll
    for (int c = 0; c < NCHUNK; c += 2) {
        step(c, t0);
        step(c + 1, t1);
    }
    __syncthreads();
    if constexpr (DB) {
        add_wave_sums(NCHUNK - 1);
        __syncthreads();
    }
    if (tid < BLK_PTS) {
        const int k = tid;
        float sum = 0.f;
        if constexpr (DB) {
            sum = psum[k];
        } else {
#pragma unroll
            for (int w = 0; w < W; w++) sum += rcw[w * BLK_PTS + k];
        }
        const int x1 = B.g1 + k / (BLK * BLK), x2 = B.g2 + (k / BLK) % BLK, x3 = B.g3 + k % BLK;
        if (x1 < n1 && x2 < n2 && x3 < n3) atomicAdd(&rho[((int64_t)x1 * n2 + x2) * n3 + x3], (double)sum);
    }
}

// ---- bf16x3 density, chunks loaded straight into LDS -------------------------
// The 16-wave density with the chunk planes brought in by LDS-DMA loads
// (global_load_lds_dwordx4: lane l of a wave writes its 16 bytes at lane * 16
// of a 1 KB fragment), double-buffered: while chunk c is multiplied out of
// one buffer, the loads of chunk c + 1 fill the other, with no prefetch
// registers and one barrier per chunk.  The two buffers are distinct LDS
// variables, so the compiler can tell the loads in flight into one from the
// reads of the other.  The registers this frees hold one accumulator per
// 16-row group: the products of the four groups are issued term by term
// across the groups (each accumulator sees the terms in the order of
// k_rblk_bf, so the sums are the same bits) and their chains no longer wait
// on one another, and the dots of a 16-point group follow its products.
// Limits: R <= DMA_ROWS (two fragments per wave and plane).
constexpr int DMA_ROWS = 512;
// fragments per wave and plane of the 16-wave density kernels
inline int nfrag_rho(int rows) { return (rows / 16 + RB_WAVES - 1) / RB_WAVES; }
constexpr int DMA_PLANE = DMA_ROWS * CHUNK;  // bf16 elements of one plane of one buffer

struct DmaLds {
    uint16_t *hi0, *lo0, *hi1, *lo1;  // chunk buffers 0 and 1 (point-fragment order)
    float *wsum;                      // [chunk % 2][wave][32 points]
    float *psum;                      // [512 points]
    int16_t *rinfo;
    SlotPair *sp;
};

__device__ __forceinline__ void dma_frag(const GLOBAL_AS uint16_t *src, uint16_t *dst) {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_global_load_lds((const GLOBAL_AS void *)src, (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
#endif
}

template <int NF>
__device__ __forceinline__ void rblk_item_dma(const HItem it, const BlockArgs &A, const DmaLds &L) {
    constexpr int W = RB_WAVES;
    const BlockDev B = as_global(A.blocks)[it.block];
    const GLOBAL_AS uint16_t *__restrict__ thi = as_global(A.thi);
    const GLOBAL_AS uint16_t *__restrict__ tlo = as_global(A.tlo);
    const GLOBAL_AS int16_t *__restrict__ rowinfo = as_global(A.rowinfo);
    const GLOBAL_AS SlotPair *__restrict__ spair = as_global(A.spair);
    const GLOBAL_AS float *__restrict__ Deff = as_global(A.Deff);
    double *__restrict__ rho = A.rho;
    const int n1 = A.n1, n2 = A.n2, n3 = A.n3;
    const int R = B.rows, S = R / ST, G = R / 16, G32 = R / 32;
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63, l16 = lane & 15,
              q = lane >> 4;
    for (int x = tid; x < R; x += 64 * W) L.rinfo[x] = rowinfo[B.row_off + x];
    for (int x = tid; x < B.nslot * B.nslot; x += 64 * W) L.sp[x] = spair[B.sp_off + x];
    // chunk c's fragments g = wave + 16 f of both planes into buffer c % 2
    auto load_chunk = [&](int c, uint16_t *bh, uint16_t *bl) {
        const GLOBAL_AS uint16_t *sh = thi + B.tile_off + (int64_t)c * G * FRAG;
        const GLOBAL_AS uint16_t *sl = tlo + B.tile_off + (int64_t)c * G * FRAG;
#pragma unroll
        for (int f = 0; f < NF; f++) {
            const int g = wave + f * W;
            if (g < G) {  // uniform across the wave
                dma_frag(sh + (g * 64 + lane) * 8, bh + g * FRAG);
                dma_frag(sl + (g * 64 + lane) * 8, bl + g * FRAG);
            }
        }
    };
    load_chunk(0, L.hi0, L.lo0);
    TouchN<64 * W> t0 = touch_chunk_n<64 * W>(thi, tlo, B.tile_off + (int64_t)G * FRAG, R, tid);
    TouchN<64 * W> t1 = touch_chunk_n<64 * W>(thi, tlo, B.tile_off + (int64_t)2 * G * FRAG, R, tid);
    __syncthreads();  // the row descriptors and slot pairs
    const bool has = wave < it.nst;
    int si = 0, sj = 0;
    if (has) st_decode(it.st0 + wave, S, si, sj);
    bf16x8 wh[4][2], wl[4][2];
    gather_w16(Deff, L.rinfo, L.sp, B.nslot, si, sj, has, l16, q, wh, wl);
    auto add_wave_sums = [&](int c) {  // the 16 wave sums of chunk c's points, in wave order
        if (wave == c % W && lane < CHUNK) {
            const float *src = L.wsum + (c & 1) * W * CHUNK;
            float sum = 0.f;
#pragma unroll
            for (int w0 = 0; w0 < W; w0 += 4) {
                float v[4];
#pragma unroll
                for (int k = 0; k < 4; k++) v[k] = src[(w0 + k) * CHUNK + lane];
#pragma unroll
                for (int k = 0; k < 4; k++) sum += v[k];
                __builtin_amdgcn_sched_barrier(0);
            }
            L.psum[c * CHUNK + lane] = sum;
        }
    };
    auto step = [&](int c, const uint16_t *BHs, const uint16_t *BLs, uint16_t *NH, uint16_t *NL, TouchN<64 * W> &slot) {
        // this wave's loads of chunk c (and its touches) are complete, then
        // every wave's: chunk c is in LDS and chunk c - 1's buffer is free
        __builtin_amdgcn_s_waitcnt(0x3f70);  // vmcnt(0)
        release_n(slot);
        __syncthreads();
        if (c > 0) add_wave_sums(c - 1);
        if (c + 1 < NCHUNK) load_chunk(c + 1, NH, NL);
        // chunk c + 2 into L2 one step before its loads (further ahead, L2 thrashes)
        if (c + 2 < NCHUNK) slot = touch_chunk_n<64 * W>(thi, tlo, B.tile_off + (int64_t)(c + 2) * G * FRAG, R, tid);
        const u16x8 *BH = reinterpret_cast<const u16x8 *>(BHs);
        const u16x8 *BL = reinterpret_cast<const u16x8 *>(BLs);
#pragma unroll 1
        for (int pt = 0; pt < 2; pt++) {
            float part = 0.f;
            if (has) {  // a wave without a super-tile only joins the sums and barriers
            f32x4 acc[4];
#pragma unroll
            for (int rt = 0; rt < 4; rt++) acc[rt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 2; ks++) {
                const int fi = pt * G32 + sj * 2 + ks;
                const bf16x8 bh = __builtin_bit_cast(bf16x8, BH[fi * 64 + lane]);
                const bf16x8 bl = __builtin_bit_cast(bf16x8, BL[fi * 64 + lane]);
#pragma unroll
                for (int rt = 0; rt < 4; rt++) acc[rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wl[rt][ks], bh, acc[rt], 0, 0, 0);
#pragma unroll
                for (int rt = 0; rt < 4; rt++) acc[rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh[rt][ks], bl, acc[rt], 0, 0, 0);
#pragma unroll
                for (int rt = 0; rt < 4; rt++) acc[rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh[rt][ks], bh, acc[rt], 0, 0, 0);
            }
            // acc[rt][v] = Q[row si*64 + 16 rt + 4 q + v][point 16 pt + l16]; Phi at the
            // same four rows and point: 8 contiguous bytes of a point fragment
#pragma unroll
            for (int rt = 0; rt < 4; rt++) {
                const int fi = pt * G32 + si * 2 + (rt >> 1);
                const int off = (fi * 64 + (2 * (rt & 1) + (q >> 1)) * 16 + l16) * 8 + 4 * (q & 1);
                const u16x4 ph = *reinterpret_cast<const u16x4 *>(BHs + off);
                const u16x4 pl = *reinterpret_cast<const u16x4 *>(BLs + off);
#pragma unroll
                for (int v = 0; v < 4; v++) part = fmaf(bf_bits_to_f32(ph[v]) + bf_bits_to_f32(pl[v]), acc[rt][v], part);
            }
            }
            const auto s32 = __builtin_amdgcn_permlane32_swap(__float_as_uint(part), __float_as_uint(part), false, false);
            part = __uint_as_float(s32[0]) + __uint_as_float(s32[1]);
            const auto s16 = __builtin_amdgcn_permlane16_swap(__float_as_uint(part), __float_as_uint(part), false, false);
            part = __uint_as_float(s16[0]) + __uint_as_float(s16[1]);
            if (q == 0) L.wsum[(c & 1) * W * CHUNK + wave * CHUNK + pt * 16 + l16] = part;
        }
    };
    // fully unrolled: the buffers alternate as distinct variables
#pragma unroll
    for (int c = 0; c < NCHUNK; c += 2) {
        step(c, L.hi0, L.lo0, L.hi1, L.lo1, t0);
        step(c + 1, L.hi1, L.lo1, L.hi0, L.lo0, t1);
    }
    __syncthreads();
    add_wave_sums(NCHUNK - 1);
    __syncthreads();
    if (tid < BLK_PTS) {
        const int k = tid;
        const float sum = L.psum[k];
        const int x1 = B.g1 + k / (BLK * BLK), x2 = B.g2 + (k / BLK) % BLK, x3 = B.g3 + k % BLK;
        if (x1 < n1 && x2 < n2 && x3 < n3) atomicAdd(&rho[((int64_t)x1 * n2 + x2) * n3 + x3], (double)sum);
    }
}

template <int NF, int MODE>
__global__ __launch_bounds__(64 * RB_WAVES) void k_rblk_dma(const HItem *__restrict__ items, int32_t nitems,
                                                           const BlockArgs *__restrict__ args, BlockArgs one,
                                                           int32_t *__restrict__ queue) {
    __shared__ __align__(16) uint16_t s_hi0[DMA_PLANE], s_lo0[DMA_PLANE], s_hi1[DMA_PLANE], s_lo1[DMA_PLANE];
    __shared__ __align__(16) float s_wsum[2 * RB_WAVES * CHUNK], s_psum[BLK_PTS];
    __shared__ __align__(16) int16_t s_rinfo[MAX_ROWS];
    __shared__ __align__(16) SlotPair s_sp[MAX_SLOT_PAIRS];
    const DmaLds L{s_hi0, s_lo0, s_hi1, s_lo1, s_wsum, s_psum, s_rinfo, s_sp};
    run_items<MODE>(items, nitems, args, one, queue,
                    [&](const HItem &it, const BlockArgs &A) { rblk_item_dma<NF>(it, A, L); });
}

bool rho_dma_fits(int rows) { return rows <= DMA_ROWS; }

hipError_t launch_rblk_dma(hipStream_t stream, int rows, const HItem *items, int32_t nitems, const BlockArgs *args,
                           const BlockArgs &one) {
    if (!rho_dma_fits(rows)) return hipErrorInvalidValue;
    const int mode = launch_mode(args, nullptr);
    void (*kern)(const HItem *, int32_t, const BlockArgs *, BlockArgs, int32_t *) = nullptr;
    if (nfrag_rho(rows) == 1) kern = mode == LAUNCH_ONE ? k_rblk_dma<1, LAUNCH_ONE> : k_rblk_dma<1, LAUNCH_BATCH>;
    else kern = mode == LAUNCH_ONE ? k_rblk_dma<2, LAUNCH_ONE> : k_rblk_dma<2, LAUNCH_BATCH>;
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(kern, dim3((unsigned)nitems), dim3(64 * RB_WAVES), 0, stream, items, nitems, args, one, nullptr);
    return hipGetLastError();
}

// ---- bf16x3 density, two workgroups per CU -----------------------------------
// The density of k_rblk_bf in items of at most RH_WAVES super-tiles (one per
// wave, 8 waves, at most 128 VGPRs) with 68 KB of LDS, so that two
// workgroups share a CU: their barriers are independent, and the multiplies
// of one run while the other waits for its tiles or gathers its W.  The point
// fragments of a chunk are two independent 16-point halves (each is all that
// the products and dots of its points read), so the tiles stream through two
// half-chunk buffers (distinct LDS variables) filled by LDS-DMA loads: half h
// is multiplied out of one while half h + 1 lands in the other, one barrier
// per half.  A block's items follow one another on one XCD, whose L2 then
// serves the second its tiles.  Each point's sum over a wave's super-tile and
// over the waves (in wave order, by one wave after the next barrier) as in
// k_rblk_bf, and one FP64 atomic per item and point.  Limits: R <= RH_ROWS
// (the plan gives the larger blocks to the 16-wave kernel).
constexpr int RH_HALF = RH_ROWS * 16;  // bf16 elements of one plane of one half-chunk buffer

struct RhLds {
    uint16_t *hi0, *lo0, *hi1, *lo1;  // half-chunk buffers 0 and 1
    float *wsum;                      // [half % 2][wave][16 points]
    float *psum;                      // [512 points]
    int16_t *rinfo;
    SlotPair *sp;
};

__device__ __forceinline__ void rblk_item_w8(const HItem it, const BlockArgs &A, const RhLds &L) {
    constexpr int W = RH_WAVES, NH = 2 * NCHUNK, HP = CHUNK / 2;
    const BlockDev B = as_global(A.blocks)[it.block];
    const GLOBAL_AS uint16_t *__restrict__ thi = as_global(A.thi);
    const GLOBAL_AS uint16_t *__restrict__ tlo = as_global(A.tlo);
    const GLOBAL_AS int16_t *__restrict__ rowinfo = as_global(A.rowinfo);
    const GLOBAL_AS SlotPair *__restrict__ spair = as_global(A.spair);
    const GLOBAL_AS float *__restrict__ Deff = as_global(A.Deff);
    double *__restrict__ rho = A.rho;
    const int n1 = A.n1, n2 = A.n2, n3 = A.n3;
    const int R = B.rows, S = R / ST, G32 = R / 32;
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63, l16 = lane & 15,
              q = lane >> 4;
    for (int x = tid; x < R; x += 64 * W) L.rinfo[x] = rowinfo[B.row_off + x];
    for (int x = tid; x < B.nslot * B.nslot; x += 64 * W) L.sp[x] = spair[B.sp_off + x];
    // half h (chunk h / 2, points 16 (h % 2) ..): fragments g = wave + 8 f < G32 of both planes
    auto load_half = [&](int h, uint16_t *bh, uint16_t *bl) {
        const GLOBAL_AS uint16_t *sh = thi + B.tile_off + (int64_t)h * G32 * FRAG;
        const GLOBAL_AS uint16_t *sl = tlo + B.tile_off + (int64_t)h * G32 * FRAG;
#pragma unroll
        for (int f = 0; f < (RH_ROWS / 32 + W - 1) / W; f++) {
            const int g = wave + f * W;
            if (g < G32) {  // uniform across the wave
                dma_frag(sh + (g * 64 + lane) * 8, bh + g * FRAG);
                dma_frag(sl + (g * 64 + lane) * 8, bl + g * FRAG);
            }
        }
    };
    load_half(0, L.hi0, L.lo0);
    __syncthreads();  // the row descriptors and slot pairs
    const bool has = wave < it.nst;
    int si = 0, sj = 0;
    if (has) st_decode(it.st0 + wave, S, si, sj);
    bf16x8 wh[4][2], wl[4][2];
    gather_w16(Deff, L.rinfo, L.sp, B.nslot, si, sj, has, l16, q, wh, wl);
    auto add_wave_sums = [&](int h) {  // the 8 wave sums of half h's points, in wave order
        if (wave == h % W && lane < HP) {
            const float *src = L.wsum + (h & 1) * W * HP;
            float sum = 0.f;
#pragma unroll
            for (int w = 0; w < W; w++) sum += src[w * HP + lane];
            L.psum[h * HP + lane] = sum;
        }
    };
    auto step = [&](int h, const uint16_t *BHs, const uint16_t *BLs, uint16_t *NH_, uint16_t *NL_) {
        // this wave's loads of half h are complete, then every wave's: half h
        // is in LDS and the other buffer (half h - 1) is free
        __builtin_amdgcn_s_waitcnt(0x3f70);  // vmcnt(0)
        __syncthreads();
        if (h > 0) add_wave_sums(h - 1);
        if (h + 1 < NH) load_half(h + 1, NH_, NL_);
        const u16x8 *BH = reinterpret_cast<const u16x8 *>(BHs);
        const u16x8 *BL = reinterpret_cast<const u16x8 *>(BLs);
        float part = 0.f;
        if (has) {
            bf16x8 bh[2], bl[2];
#pragma unroll
            for (int ks = 0; ks < 2; ks++) {
                bh[ks] = __builtin_bit_cast(bf16x8, BH[(sj * 2 + ks) * 64 + lane]);
                bl[ks] = __builtin_bit_cast(bf16x8, BL[(sj * 2 + ks) * 64 + lane]);
            }
#pragma unroll
            for (int rt = 0; rt < 4; rt++) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int ks = 0; ks < 2; ks++) {
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wl[rt][ks], bh[ks], acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh[rt][ks], bl[ks], acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh[rt][ks], bh[ks], acc, 0, 0, 0);
                }
                // acc[v] = Q[row si*64 + 16 rt + 4 q + v][point l16 of the half]
                const int fi = si * 2 + (rt >> 1);
                const int off = (fi * 64 + (2 * (rt & 1) + (q >> 1)) * 16 + l16) * 8 + 4 * (q & 1);
                const u16x4 ph = *reinterpret_cast<const u16x4 *>(BHs + off);
                const u16x4 pl = *reinterpret_cast<const u16x4 *>(BLs + off);
#pragma unroll
                for (int v = 0; v < 4; v++) part = fmaf(bf_bits_to_f32(ph[v]) + bf_bits_to_f32(pl[v]), acc[v], part);
            }
        }
        const auto s32 = __builtin_amdgcn_permlane32_swap(__float_as_uint(part), __float_as_uint(part), false, false);
        part = __uint_as_float(s32[0]) + __uint_as_float(s32[1]);
        const auto s16 = __builtin_amdgcn_permlane16_swap(__float_as_uint(part), __float_as_uint(part), false, false);
        part = __uint_as_float(s16[0]) + __uint_as_float(s16[1]);
        if (q == 0) L.wsum[(h & 1) * W * HP + wave * HP + l16] = part;
    };
    // fully unrolled: the buffers alternate as distinct variables
#pragma unroll
    for (int h = 0; h < NH; h += 2) {
        step(h, L.hi0, L.lo0, L.hi1, L.lo1);
        step(h + 1, L.hi1, L.lo1, L.hi0, L.lo0);
    }
    __syncthreads();
    add_wave_sums(NH - 1);
    __syncthreads();
    {
        const int k = tid;  // 512 threads, one point each
        const float sum = L.psum[k];
        const int x1 = B.g1 + k / (BLK * BLK), x2 = B.g2 + (k / BLK) % BLK, x3 = B.g3 + k % BLK;
        if (x1 < n1 && x2 < n2 && x3 < n3) atomicAdd(&rho[((int64_t)x1 * n2 + x2) * n3 + x3], (double)sum);
    }
}

template <int MODE>
__global__ __launch_bounds__(64 * RH_WAVES) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_rblk_w8(
    const HItem *__restrict__ items, int32_t nitems, const BlockArgs *__restrict__ args, BlockArgs one,
    int32_t *__restrict__ queue) {
    __shared__ __align__(16) uint16_t s_hi0[RH_HALF], s_lo0[RH_HALF], s_hi1[RH_HALF], s_lo1[RH_HALF];
    __shared__ __align__(16) float s_wsum[2 * RH_WAVES * CHUNK / 2], s_psum[BLK_PTS];
    __shared__ __align__(16) int16_t s_rinfo[RH_ROWS];
    __shared__ __align__(16) SlotPair s_sp[MAX_SLOT_PAIRS];
    const RhLds L{s_hi0, s_lo0, s_hi1, s_lo1, s_wsum, s_psum, s_rinfo, s_sp};
    run_items<MODE>(items, nitems, args, one, queue, [&](const HItem &it, const BlockArgs &A) { rblk_item_w8(it, A, L); });
}

bool rho_w8_fits(int rows) { return rows <= RH_ROWS; }

hipError_t launch_rblk_w8(hipStream_t stream, int rows, const HItem *items, int32_t nitems, const BlockArgs *args,
                          const BlockArgs &one) {
    if (!rho_w8_fits(rows)) return hipErrorInvalidValue;
    const int mode = launch_mode(args, nullptr);
    void (*kern)(const HItem *, int32_t, const BlockArgs *, BlockArgs, int32_t *) =
        mode == LAUNCH_ONE ? k_rblk_w8<LAUNCH_ONE> : k_rblk_w8<LAUNCH_BATCH>;
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(kern, dim3((unsigned)nitems), dim3(64 * RH_WAVES), 0, stream, items, nitems, args, one, nullptr);
    return hipGetLastError();
}

// ---- FP32 block density -----------------------------------------------------
// rho_k = sum_{si<=sj} Phi_si^T W[si][sj] Phi_sj as in k_rblk_bf, with FP32
// operands on the FP32 matrix cores (v_mfma_f32_16x16x4f32): one 1 024-thread
// workgroup per item, one super-tile per wave.  The wave's W super-tile is
// held as A fragments (lane l of fragment (rt, kt): W[16 rt + l % 16][4 kt +
// l / 16], 64 VGPRs), gathered from the FP32 fold as the bf16x3 kernel does;
// per 32-point chunk the block's FP32 tile (R x 32, the H kernel's tile in
// its perm32 point order) is staged in LDS, Q = W Phi_sj is formed per
// 16-point group and dotted with Phi_si at the same rows, and the four lanes
// of a point are summed with the lane swaps into per-wave LDS slots, added in
// wave order at the end of the item (one FP64 atomic per item and point).
constexpr int RF_WAVES = 16;
constexpr int RF_SUMS_OFFSET = MAX_ROWS * LROW32 * 4;
constexpr int RF_AREA_BYTES = RF_SUMS_OFFSET + RF_WAVES * BLK_PTS * 4;

// One point per lane of a 16-point group: Q = W Phi_sj at the group's
// points on the FP32 matrix cores, dotted with Phi_si; DIAG skips the
// (zero) tiles below a diagonal super-tile's diagonal.
template <bool DIAG>
__device__ __forceinline__ float rblk_f32_point(const float *__restrict__ A32, const float (&w)[4][16], int si, int sj,
                                                int q, int pcol) {
    f32x4 acc[4];
#pragma unroll
    for (int rt = 0; rt < 4; rt++) acc[rt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 16; kt++) {
        const float b = A32[(sj * ST + 4 * kt + q) * LROW32 + pcol];
#pragma unroll
        for (int rt = 0; rt < 4; rt++) {
            if (DIAG && rt > kt / 4) continue;
            acc[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(w[rt][kt], b, acc[rt], 0, 0, 0);
        }
    }
    // acc[rt][v] = Q[16 rt + 4 q + v][point 16 pt + l16]
    float part = 0.f;
#pragma unroll
    for (int rt = 0; rt < 4; rt++)
#pragma unroll
        for (int v = 0; v < 4; v++) part = fmaf(A32[(si * ST + 16 * rt + 4 * q + v) * LROW32 + pcol], acc[rt][v], part);
    return part;
}

__device__ __forceinline__ void rblk_f32_item(const HItem it, const BlockArgs &A, unsigned char *lds_raw) {
    const BlockDev B = as_global(A.blocks)[it.block];
    const GLOBAL_AS float *__restrict__ t32 = as_global(A.t32);
    const GLOBAL_AS int16_t *__restrict__ rowinfo = as_global(A.rowinfo);
    const GLOBAL_AS SlotPair *__restrict__ spair = as_global(A.spair);
    const GLOBAL_AS float *__restrict__ Deff = as_global(A.Deff);
    GLOBAL_AS double *__restrict__ rho = as_global(A.rho);
    const int n1 = A.n1, n2 = A.n2, n3 = A.n3;
    const int R = B.rows, S = R / ST;
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63, l16 = lane & 15,
              q = lane >> 4;
    float *A32 = reinterpret_cast<float *>(lds_raw);
    float *rcw = reinterpret_cast<float *>(lds_raw + RF_SUMS_OFFSET);  // [wave][512 points]
    int16_t *s_rinfo = reinterpret_cast<int16_t *>(lds_raw + RF_AREA_BYTES);
    SlotPair *s_sp = reinterpret_cast<SlotPair *>(s_rinfo + MAX_ROWS);
    for (int x = tid; x < R; x += 64 * RF_WAVES) s_rinfo[x] = rowinfo[B.row_off + x];
    for (int x = tid; x < B.nslot * B.nslot; x += 64 * RF_WAVES) s_sp[x] = spair[B.sp_off + x];
    __syncthreads();
    const bool has = wave < it.nst;
    int si = 0, sj = 0;
    if (has) st_decode(it.st0 + wave, S, si, sj);
    // W fragments: w[rt][kt] = W[16 rt + l16][4 kt + q] of super-tile (si, sj)
    // A diagonal super-tile (si == sj) holds a symmetric W: its 16x16 tiles
    // above the diagonal carry W + W^T (factor 1 of the fold D + D^T), the
    // diagonal tiles (D + D^T) / 2, and those below are zero and skipped.
    const bool diag = si == sj;
    float w[4][16];
    {
        const float keep = has ? 1.f : 0.f;
#pragma unroll
        for (int rt = 0; rt < 4; rt++) {
            const int rinf = s_rinfo[si * ST + 16 * rt + l16];
            const int ra = rinf >> 5, ri = rinf & 31;
#pragma unroll
            for (int kt = 0; kt < 16; kt++) {
                const int cinf = s_rinfo[sj * ST + 4 * kt + q];
                const int cb = cinf >> 5, cj = cinf & 31;
                const bool valid = rinf >= 0 && cinf >= 0;
                const SlotPair P = s_sp[valid ? min(ra, cb) * B.nslot + max(ra, cb) : 0];
                const bool sw = ra > cb;  // lower triangle of a diagonal super-tile: W is symmetric
                const int i = sw ? cj : ri, j = sw ? ri : cj;
                const bool ok = valid && P.hoff >= 0;
                const int ld = P.meta & 63;
                const int x1 = (P.meta >> 6) ? j * ld + i : i * ld + j;
                // D is finite, so masking by a product keeps the load unconditional
                const int ct = kt / 4;
                const float wdiag = !diag ? 1.f : rt < ct ? 1.f : rt == ct ? 0.5f : 0.f;
                w[rt][kt] = Deff[ok ? P.hoff + x1 : 0] * (ok ? wdiag * keep : 0.f);
            }
        }
    }
    for (int c = 0; c < NCHUNK; c++) {
        __syncthreads();  // every wave is done with the previous chunk
        const GLOBAL_AS float *s32 = t32 + B.tile_off + (int64_t)c * R * CHUNK;
        for (int e = tid * 4; e < R * CHUNK; e += 64 * RF_WAVES * 4) {
            const int row = e / CHUNK, col = e % CHUNK;
            *reinterpret_cast<f32x4 *>(&A32[row * LROW32 + col]) = *reinterpret_cast<const f32x4 *>(&s32[e]);
        }
        __syncthreads();
#pragma unroll 1
        for (int pt = 0; pt < 2; pt++) {
            // the lane's point: 16 pt + l16, stored at perm32 of it
            const int pcol = perm32(16 * pt + l16);
            float part = 0.f;
            if (has) part = diag ? rblk_f32_point<true>(A32, w, si, sj, q, pcol)
                                 : rblk_f32_point<false>(A32, w, si, sj, q, pcol);
            const auto s32s = __builtin_amdgcn_permlane32_swap(__float_as_uint(part), __float_as_uint(part), false, false);
            part = __uint_as_float(s32s[0]) + __uint_as_float(s32s[1]);
            const auto s16 = __builtin_amdgcn_permlane16_swap(__float_as_uint(part), __float_as_uint(part), false, false);
            part = __uint_as_float(s16[0]) + __uint_as_float(s16[1]);
            if (q == 0) rcw[wave * BLK_PTS + c * CHUNK + pt * 16 + l16] = part;
        }
    }
    __syncthreads();
    if (tid < BLK_PTS) {
        const int k = tid;
        float sum = 0.f;
#pragma unroll
        for (int w2 = 0; w2 < RF_WAVES; w2++) sum += rcw[w2 * BLK_PTS + k];
        const int x1 = B.g1 + k / (BLK * BLK), x2 = B.g2 + (k / BLK) % BLK, x3 = B.g3 + k % BLK;
        if (x1 < n1 && x2 < n2 && x3 < n3) atomicAdd(&rho[((int64_t)x1 * n2 + x2) * n3 + x3], (double)sum);
    }
}

template <int MODE>
__global__ __launch_bounds__(64 * RF_WAVES) void k_rblk_f32(const HItem *__restrict__ items, int32_t nitems,
                                                            const BlockArgs *__restrict__ args, BlockArgs one) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    run_items<MODE>(items, nitems, args, one, nullptr,
                    [&](const HItem &it, const BlockArgs &A) { rblk_f32_item(it, A, lds_raw); });
}

hipError_t launch_f32_block_kernel(bool density, hipStream_t stream, int rows, int max_slots, const HItem *items,
                                   int32_t nitems, const BlockArgs *args, const BlockArgs &one) {
    if (nitems <= 0) return hipSuccess;
    const bool batch = args != nullptr;
    const void *kern;
    size_t lds;
    if (density) {
        kern = batch ? reinterpret_cast<const void *>(k_rblk_f32<LAUNCH_BATCH>)
                     : reinterpret_cast<const void *>(k_rblk_f32<LAUNCH_ONE>);
        lds = RF_AREA_BYTES + MAX_ROWS * sizeof(int16_t) + (size_t)max_slots * max_slots * sizeof(SlotPair);
    } else {
        kern = batch ? reinterpret_cast<const void *>(k_hblk_f32<LAUNCH_BATCH>)
                     : reinterpret_cast<const void *>(k_hblk_f32<LAUNCH_ONE>);
        lds = (size_t)rows * LROW32 * sizeof(float) + CHUNK * sizeof(float);
    }
    hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const unsigned threads = density ? 64 * RF_WAVES : 64 * HB_WAVES;
    if (density) {
        if (batch) hipLaunchKernelGGL(k_rblk_f32<LAUNCH_BATCH>, dim3(nitems), dim3(threads), lds, stream, items, nitems, args, one);
        else hipLaunchKernelGGL(k_rblk_f32<LAUNCH_ONE>, dim3(nitems), dim3(threads), lds, stream, items, nitems, args, one);
    } else {
        if (batch) hipLaunchKernelGGL(k_hblk_f32<LAUNCH_BATCH>, dim3(nitems), dim3(threads), lds, stream, items, nitems, args, one);
        else hipLaunchKernelGGL(k_hblk_f32<LAUNCH_ONE>, dim3(nitems), dim3(threads), lds, stream, items, nitems, args, one);
    }
    return hipGetLastError();
}

// fragments per wave and plane of the density kernel for R rows
constexpr int MAXF_RHO = (MAX_ROWS / 16 + RB_WAVES - 1) / RB_WAVES;

// waves of the density workgroup: 16 for one structure; the batched
// (persistent) form keeps 8, whose loop state would not fit 128 VGPRs
template <int MODE>
constexpr int rb_waves() { return MODE == LAUNCH_QUEUE ? HB_WAVES : RB_WAVES; }

template <int NF, int MODE, bool DB = false>
__global__ __launch_bounds__(64 * rb_waves<MODE>()) void k_rblk_bf(const HItem *__restrict__ items, int32_t nitems,
                                                                  const BlockArgs *__restrict__ args, BlockArgs one,
                                                                  int32_t *__restrict__ queue) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    run_items<MODE>(items, nitems, args, one, queue, [&](const HItem &it, const BlockArgs &A) {
        rblk_item<NF, rb_waves<MODE>(), DB>(it, A, lds_raw);
    });
}

BlockArgs block_args(const lcpao_ctx *ctx, const double *V, double *rho) {
    const BlockPlan *P = ctx->plan;
    BlockArgs a{};
    a.blocks = P->d_blocks;
    a.hi = P->d_hi; a.lo = P->d_lo; a.thi = P->d_thi; a.tlo = P->d_tlo;
    a.rowinfo = P->d_rowinfo; a.spair = P->d_spair;
    a.V = V; a.Deff = ctx->d_deff32; a.part = P->d_part; a.rho = rho;
    a.slotno = P->d_slotno; a.t32 = P->d_t32;
    a.n1 = ctx->ngrid[0]; a.n2 = ctx->ngrid[1]; a.n3 = ctx->ngrid[2];
    return a;
}

constexpr size_t LDS_PER_CU = 160 * 1024;

bool rho_db_fits(int rows, int max_slots) {
    // up to two prefetch fragments per wave (R <= 512): with three the
    // double-buffered kernel spills
    return nfrag_rho(rows) <= 2 &&
           RB_DB_AREA_BYTES + MAX_ROWS * sizeof(int16_t) + (size_t)max_slots * max_slots * sizeof(SlotPair) <=
               LDS_PER_CU;
}

size_t block_kernel_lds(bool density, int rows, int max_slots, bool persistent, bool db) {
    // density: chunk planes and point sums (the 8-wave form: W images first;
    // the double-buffered form: two plane buffers and the chunk sums), then
    // the block's row descriptors and slot pairs
    if (density)
        return (persistent ? RB_IMAGE_BYTES : db ? RB_DB_AREA_BYTES : RB_AREA_BYTES) + MAX_ROWS * sizeof(int16_t) +
               (size_t)max_slots * max_slots * sizeof(SlotPair);
    return (size_t)4 * rows * CHUNK * sizeof(uint16_t) + BLK_PTS * sizeof(float);
}

hipError_t launch_block_kernel(bool density, hipStream_t stream, int rows, int max_slots, const HItem *items,
                               int32_t nitems, const BlockArgs *args, const BlockArgs &one, int32_t *queue,
                               unsigned grid, bool db) {
    const int mode = launch_mode(args, queue);
    const bool persistent = mode == LAUNCH_QUEUE;
    if (db && (!density || persistent || !rho_db_fits(rows, max_slots))) return hipErrorInvalidValue;
    const size_t lds = block_kernel_lds(density, rows, max_slots, persistent, db);
    void (*kern)(const HItem *, int32_t, const BlockArgs *, BlockArgs, int32_t *) = nullptr;
#define BY_MODE(K, NF)                                                                             \
    kern = mode == LAUNCH_ONE ? K<NF, LAUNCH_ONE> : mode == LAUNCH_BATCH ? K<NF, LAUNCH_BATCH> \
                                                                         : K<NF, LAUNCH_QUEUE>
#define BY_MODE_DB(K, NF) kern = mode == LAUNCH_ONE ? K<NF, LAUNCH_ONE, true> : K<NF, LAUNCH_BATCH, true>
    if (density && !persistent && db) {
        if (nfrag_rho(rows) == 1) BY_MODE_DB(k_rblk_bf, 1);
        else BY_MODE_DB(k_rblk_bf, 2);
    } else if (density && !persistent) {
        static_assert(MAXF_RHO <= 3, "density prefetch depths instantiated: 1 .. 3");
        switch (nfrag_rho(rows)) {
            case 1: BY_MODE(k_rblk_bf, 1); break;
            case 2: BY_MODE(k_rblk_bf, 2); break;
            default: BY_MODE(k_rblk_bf, 3); break;
        }
    } else if (density) {
        switch (nfrag(rows)) {
            case 1: kern = k_rblk_bf<1, LAUNCH_QUEUE>; break;
            case 2: kern = k_rblk_bf<2, LAUNCH_QUEUE>; break;
            case 3: kern = k_rblk_bf<3, LAUNCH_QUEUE>; break;
            case 4: kern = k_rblk_bf<4, LAUNCH_QUEUE>; break;
            default: kern = k_rblk_bf<MAXF, LAUNCH_QUEUE>; break;
        }
    } else {
        switch (nfrag(rows)) {
            case 1: BY_MODE(k_hblk_bf, 1); break;
            case 2: BY_MODE(k_hblk_bf, 2); break;
            case 3: BY_MODE(k_hblk_bf, 3); break;
            case 4: BY_MODE(k_hblk_bf, 4); break;
            default: BY_MODE(k_hblk_bf, MAXF); break;
        }
    }
#undef BY_MODE
#undef BY_MODE_DB
    const int threads = 64 * (density && !persistent ? RB_WAVES : HB_WAVES);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
    if (nitems <= 0) return hipSuccess;
    if (grid == 0) {  // persistent: as many workgroups as can be resident at once
        int dev = 0, ncu = 0, per_cu = 0;
        if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
        if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
        if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(kern),
                                                              threads, lds)) != hipSuccess)
            return e;
        grid = (unsigned)std::min<int64_t>(nitems, (int64_t)ncu * std::max(1, per_cu));
    }
    hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), lds, stream, items, nitems, args, one, queue);
    return hipGetLastError();
}

int split_density_f32_enqueue(lcpao_ctx *ctx, int nspin, const double *const *d_DM, int64_t dm_offset,
                              double *const *d_rho) {
    BlockPlan *P = ctx->plan;
    if (!P || !P->f32_ready) { ctx->err = "FP32 block tiles are not built"; return LCPAO_ESTATE; }
    for (int s = 0; s < nspin; s++) {
        if (!d_DM[s] || !d_rho[s]) { ctx->err = "null spin pointer"; return LCPAO_EINVAL; }
        if (ctx->ncanon > 0) {
            KTimer kt(ctx, KID_DEFF);
            launch_deff_f32(ctx, d_DM[s] + dm_offset, 1);
        }
        if (ctx->nn > 0)
            hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((ctx->nn + 255) / 256)), dim3(256), 0, ctx->stream, ctx->nn,
                               d_rho[s], 0.0);
        if (!P->items.empty()) {
            KTimer kt(ctx, KID_RHO);
            const unsigned ni = (unsigned)P->items.size();
            SPLIT_TRY(ctx, launch_f32_block_kernel(true, ctx->stream, P->max_rows, P->max_slots, P->d_items,
                                                   (int32_t)ni, nullptr, block_args(ctx, nullptr, d_rho[s])));
        }
        SPLIT_TRY(ctx, hipGetLastError());
    }
    return LCPAO_OK;
}

int split_density(lcpao_ctx *ctx, int nspin, const double *const *d_DM, int64_t dm_offset, double *const *d_rho) {
    int rc = split_tables(ctx, LCPAO_BF16X3);
    if (rc) return rc;
    BlockPlan *P = ctx->plan;
    for (int s = 0; s < nspin; s++) {
        if (!d_DM[s] || !d_rho[s]) { ctx->err = "null spin pointer"; return LCPAO_EINVAL; }
        const double *D = d_DM[s] + dm_offset;
        if (ctx->ncanon > 0) {
            KTimer kt(ctx, KID_DEFF);
            launch_deff_f32(ctx, D, 1);
        }
        if (ctx->nn > 0)
            hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((ctx->nn + 255) / 256)), dim3(256), 0, ctx->stream, ctx->nn,
                               d_rho[s], 0.0);
        if (!P->items.empty()) {
            const unsigned ni = (unsigned)P->items.size();
            const BlockArgs one = block_args(ctx, nullptr, d_rho[s]);
            // candidates: 0 = single-buffered chunk planes, 1 = double-buffered
            // (when its LDS fits the plan's largest slot count), 2 = chunks
            // loaded straight into LDS (R <= DMA_ROWS), 3 = 8-wave items, two
            // workgroups per CU (blocks of at most RH_ROWS rows; the larger
            // ones in a second launch of the 16-wave kernel)
            auto fits = [&](int v) {
                return v == 0 || v == 3 || (v == 1 && rho_db_fits(P->max_rows, P->max_slots)) ||
                       (v == 2 && rho_dma_fits(P->max_rows));
            };
            auto launch = [&](int v) {
                if (v == 2) return launch_rblk_dma(ctx->stream, P->max_rows, P->d_items, (int32_t)ni, nullptr, one);
                if (v == 3) {
                    hipError_t e = launch_rblk_w8(ctx->stream, std::min(P->max_rows, RH_ROWS), P->d_items8,
                                                  (int32_t)P->items8.size(), nullptr, one);
                    if (e == hipSuccess && !P->items_big.empty())
                        e = launch_block_kernel(true, ctx->stream, P->max_rows_big, P->max_slots, P->d_items_big,
                                                (int32_t)P->items_big.size(), nullptr, one, nullptr,
                                                (unsigned)P->items_big.size());
                    return e;
                }
                return launch_block_kernel(true, ctx->stream, P->max_rows, P->max_slots, P->d_items, (int32_t)ni,
                                           nullptr, one, nullptr, ni, v == 1);
            };
            if (ctx->variant[TUNE_RHOBF] < 0) {
                // the candidates add into rho, which is cleared again before
                // the real launch
                SPLIT_TRY(ctx, tune_block_kernel(ctx, TUNE_RHOBF, launch, fits));
                if (ctx->nn > 0)
                    hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((ctx->nn + 255) / 256)), dim3(256), 0, ctx->stream,
                                       ctx->nn, d_rho[s], 0.0);
            }
            KTimer kt(ctx, KID_RHO);
            SPLIT_TRY(ctx, launch(fits(ctx->variant[TUNE_RHOBF]) ? ctx->variant[TUNE_RHOBF] : 0));
        }
        SPLIT_TRY(ctx, hipGetLastError());
    }
    SPLIT_TRY(ctx, hipStreamSynchronize(ctx->stream));
    collect_timings(ctx);
    return LCPAO_OK;
}

// ---- FP64 block density -----------------------------------------------------
// rho(p) = sum over the slot pairs x <= y of a sub-block of phi_x(p)^T W_xy
// phi_y(p), W_xy the FP64 fold of the pair (D_P + D_P'^T; an on-site pair's
// own block), in exact FP64 vector FMAs.  One 1 024-thread workgroup per
// 4x4x4 sub-block of an 8^3 block, lane = grid point:
//  1. the table column of every (slot, point) is read from the plan's point
//     map into LDS; wave ballots give each slot's presence mask;
//  2. one wave lists the slot pairs that are computed pairs and share a
//     point of the sub-block (slot-major, in order) and the LDS row of every
//     present slot's orbitals;
//  3. the columns of all present slots are staged once into LDS as
//     [orbital row][64 points] FP64 (row stride 65 doubles: the staging
//     writes, one point's consecutive orbitals per lane group, and the
//     reads, one row per wave, are both free of bank conflicts), reading
//     each table column as one contiguous run;
//  4. wave w takes the pairs w, w + 16, ...: phi_x and phi_y come from LDS,
//     W is wave-uniform and read through the scalar path (SGPR operands);
//  5. the sixteen waves' sums of every point are added in wave order and
//     the point is written once: no atomics, deterministic.
// The tables are read once per launch; the lanes compute at the pairs'
// overlap points plus the absent points of partly covered sub-blocks (1.44x
// the exact work on GaAs216).  What bounds it is the W reads: each W block
// feeds only 64 points per wave, so the scalar loads (about 49 per pair)
// mostly miss the scalar cache and the waves wait on them; sixteen waves (4
// per SIMD) hide enough of that latency to match the pair-list kernels
// (12.2 ms on GaAs216, against 18.3 ms with eight waves).  A sub-block with
// more orbital rows than the launch's LDS holds reads its columns from the
// tables instead (same sums).  Workgroups are dealt XCD-aware: workgroup w
// runs on XCD w % 8, which takes a contiguous range of sub-blocks, so
// neighbouring sub-blocks -- which read the same pair blocks of W -- meet in
// one L2.
constexpr int R64_WAVES = 16;
constexpr int R64_LD = 65;       // LDS row stride of the staged columns (doubles)
constexpr int R64_UNROLL = 8;    // staging loads in flight per thread

struct Rho64Args {
    const BlockDev *blocks;
    const int32_t *sph_atom, *sph_map;
    const int16_t *slotno;
    const SlotPair *spair;
    const int64_t *tab_off;
    const double *tab, *Deff;
    double *rho;
    int64_t nsub;              // 8 x blocks
    int32_t n1, n2, n3;
    int32_t rows_cap;          // orbital rows the staged columns hold
    int32_t max_slots;
    int32_t pad;
};

// LDS layout of the FP64 block density (bytes)
struct R64Lds {
    int phi, tm, mask, roff, rslot, pairs, part, ctl, total;
};
__host__ __device__ inline R64Lds r64_lds(int rows_cap, int max_slots) {
    R64Lds L;
    L.phi = 0;
    L.tm = L.phi + rows_cap * R64_LD * 8;
    L.mask = L.tm + max_slots * 64 * 4;
    L.mask = (L.mask + 7) / 8 * 8;
    L.roff = L.mask + max_slots * 8;
    L.rslot = L.roff + max_slots * 4;
    L.pairs = L.rslot + rows_cap * 4;
    L.part = L.pairs + max_slots * (max_slots + 1) / 2 * 4;
    L.part = (L.part + 7) / 8 * 8;
    L.ctl = L.part + R64_WAVES * 64 * 8;
    L.total = L.ctl + 16;
    return L;
}

// a^T M b, M row-major NA x NB (wave-uniform)
template <int NA, int NB>
__device__ __forceinline__ double quad64(const GLOBAL_AS double *__restrict__ M, const double (&a)[NA],
                                         const double (&b)[NB]) {
    double z = 0.0;
#pragma unroll
    for (int i = 0; i < NA; i++) {
        double y = 0.0;
#pragma unroll
        for (int j = 0; j < NB; j++) y = fma(M[i * NB + j], b[j], y);
        z = fma(a[i], y, z);
    }
    return z;
}

// where the lane's orbital values of a slot come from: staged LDS rows, or
// the table column (absent point: zeros)
struct Col64 {
    const double *lds;               // staged: row 0 of the slot, stride R64_LD
    const GLOBAL_AS double *tab;     // else: the slot's table
    int t;                           // table column (-1: absent)
};
template <int N, bool STAGED>
__device__ __forceinline__ void col64(const Col64 &c, int lane, double (&f)[N]) {
    if constexpr (STAGED) {
#pragma unroll
        for (int i = 0; i < N; i++) f[i] = c.lds[i * R64_LD + lane];
    } else {
        const GLOBAL_AS double *p = c.tab + (int64_t)(c.t < 0 ? 0 : c.t) * N;
#pragma unroll
        for (int i = 0; i < N; i++) {
            const double v = p[i];
            f[i] = c.t >= 0 ? v : 0.0;
        }
    }
}
template <bool STAGED>
__device__ __forceinline__ double col64_at(const Col64 &c, int lane, int i, int n) {
    if constexpr (STAGED) return c.lds[i * R64_LD + lane];
    else return c.t >= 0 ? c.tab[(int64_t)c.t * n + i] : 0.0;
}

template <int NX, int NY, bool STAGED>
__device__ __forceinline__ double pair64(const GLOBAL_AS double *__restrict__ M, bool tr, bool self, const Col64 &cx,
                                         const Col64 &cy, int lane) {
    double fx[NX];
    col64<NX, STAGED>(cx, lane, fx);
    if constexpr (NX == NY) {
        if (self) return quad64<NX, NX>(M, fx, fx);
    }
    double fy[NY];
    col64<NY, STAGED>(cy, lane, fy);
    return tr ? quad64<NY, NX>(M, fy, fx) : quad64<NX, NY>(M, fx, fy);
}

// any orbital counts (row-by-row loops)
template <bool STAGED>
__device__ double pair64_rt(const GLOBAL_AS double *__restrict__ M, bool tr, const Col64 &cx, int nx,
                            const Col64 &cy, int ny, int lane) {
    const Col64 &ca = tr ? cy : cx, &cb = tr ? cx : cy;
    const int na = tr ? ny : nx, nb = tr ? nx : ny;
    double z = 0.0;
    for (int i = 0; i < na; i++) {
        double y = 0.0;
        for (int j = 0; j < nb; j++) y = fma(M[i * nb + j], col64_at<STAGED>(cb, lane, j, nb), y);
        z = fma(col64_at<STAGED>(ca, lane, i, na), y, z);
    }
    return z;
}

template <bool STAGED>
__device__ double rho64_pairs(const Rho64Args &A, const BlockDev &B, unsigned char *lds, const R64Lds &L, int lane,
                              int wave, int k) {
    const int32_t *tm = reinterpret_cast<const int32_t *>(lds + L.tm);
    const int32_t *roff = reinterpret_cast<const int32_t *>(lds + L.roff);
    const int32_t *pairs = reinterpret_cast<const int32_t *>(lds + L.pairs);
    const double *phi = reinterpret_cast<const double *>(lds + L.phi);
    const int np = reinterpret_cast<const int32_t *>(lds + L.ctl)[0];
    const int ns = B.nslot;
    const GLOBAL_AS double *Deff = as_global(A.Deff);
    const GLOBAL_AS SlotPair *sp = as_global(A.spair) + B.sp_off;
    const GLOBAL_AS int16_t *sno = as_global(A.slotno) + B.slot_off;
    double z = 0.0;
    for (int q = wave; q < np; q += R64_WAVES) {
        const int xy = __builtin_amdgcn_readfirstlane(pairs[q]);
        const int x = xy >> 16, y = xy & 0xFFFF;
        const SlotPair P = sp[x * ns + y];
        const GLOBAL_AS double *M = Deff + __builtin_amdgcn_readfirstlane(P.hoff);
        const bool tr = (__builtin_amdgcn_readfirstlane(P.meta) >> 6) != 0, self = x == y;
        const int nx = __builtin_amdgcn_readfirstlane(sno[x]), ny = __builtin_amdgcn_readfirstlane(sno[y]);
        Col64 cx, cy;
        if constexpr (STAGED) {
            cx.lds = phi + (int64_t)__builtin_amdgcn_readfirstlane(roff[x]) * R64_LD;
            cy.lds = phi + (int64_t)__builtin_amdgcn_readfirstlane(roff[y]) * R64_LD;
        } else {
            const GLOBAL_AS int32_t *sa = as_global(A.sph_atom) + B.slot_off;
            cx.tab = as_global(A.tab) + as_global(A.tab_off)[__builtin_amdgcn_readfirstlane(sa[x])];
            cy.tab = as_global(A.tab) + as_global(A.tab_off)[__builtin_amdgcn_readfirstlane(sa[y])];
            cx.t = tm[x * 64 + lane];
            cy.t = tm[y * 64 + lane];
        }
#define R64_CASE(NX, NY)                                                 \
    if (nx == NX && ny == NY) {                                          \
        z += pair64<NX, NY, STAGED>(M, tr, self, cx, cy, lane);          \
        continue;                                                        \
    }
        NO_CASES(R64_CASE)
#undef R64_CASE
        z += pair64_rt<STAGED>(M, tr, cx, nx, cy, ny, lane);
    }
    return z;
}

__global__ __launch_bounds__(64 * R64_WAVES) void k_rblk64(Rho64Args A) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int64_t per = (A.nsub + 7) / 8;
    const int64_t j = blockIdx.x / 8, s = (int64_t)(blockIdx.x % 8) * per + j;
    if (j >= per || s >= A.nsub) return;
    const BlockDev B = as_global(A.blocks)[s / 8];
    const int sc = (int)(s % 8);
    const int ns = B.nslot;
    const R64Lds L = r64_lds(A.rows_cap, A.max_slots);
    int32_t *tm = reinterpret_cast<int32_t *>(lds + L.tm);
    uint64_t *mask = reinterpret_cast<uint64_t *>(lds + L.mask);
    int32_t *roff = reinterpret_cast<int32_t *>(lds + L.roff);
    int32_t *rslot = reinterpret_cast<int32_t *>(lds + L.rslot);
    int32_t *pairs = reinterpret_cast<int32_t *>(lds + L.pairs);
    double *part = reinterpret_cast<double *>(lds + L.part);
    int32_t *ctl = reinterpret_cast<int32_t *>(lds + L.ctl);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    // the lane's point of the sub-block and its index in the block
    const int p1 = 4 * (sc >> 2) + (lane >> 4), p2 = 4 * ((sc >> 1) & 1) + ((lane >> 2) & 3), p3 = 4 * (sc & 1) + (lane & 3);
    const int k = (p1 * BLK + p2) * BLK + p3;
    // 1. table columns and presence masks
    const GLOBAL_AS int32_t *map = as_global(A.sph_map) + (int64_t)B.slot_off * BLK_PTS;
    for (int x = wave; x < ns; x += R64_WAVES) {
        const int t = map[(int64_t)x * BLK_PTS + k];
        tm[x * 64 + lane] = t;
        const uint64_t m = __ballot(t >= 0);
        if (lane == 0) mask[x] = m;
    }
    __syncthreads();
    // 2. pair list and LDS rows (one wave, in slot order)
    const GLOBAL_AS int16_t *sno = as_global(A.slotno) + B.slot_off;
    if (wave == 0) {
        int rows = 0, np = 0;
        for (int x = 0; x < ns; x++) {
            const uint64_t mx = mask[x];
            if (!mx) continue;
            const int nx = sno[x];
            if (lane == 0) roff[x] = rows;
            if (rows + nx <= A.rows_cap)
                for (int i = lane; i < nx; i += 64) rslot[rows + i] = x << 16 | i;
            rows += nx;
            for (int y0 = x; y0 < ns; y0 += 64) {
                const int y = y0 + lane;
                const bool act = y < ns && (mask[y] & mx) != 0 && as_global(A.spair)[B.sp_off + (int64_t)x * ns + y].hoff >= 0;
                const uint64_t bal = __ballot(act);
                if (act) pairs[np + __popcll(bal & ((1ull << lane) - 1))] = x << 16 | y;
                np += __popcll(bal);
            }
        }
        if (lane == 0) { ctl[0] = np; ctl[1] = rows; }
    }
    __syncthreads();
    const int rows = ctl[1];
    double z;
    if (rows <= A.rows_cap) {
        // 3. stage the present slots' columns: element e = point * rows + row,
        // so a lane group reads one point's orbitals of a slot contiguously
        double *phi = reinterpret_cast<double *>(lds + L.phi);
        const GLOBAL_AS int32_t *sa = as_global(A.sph_atom) + B.slot_off;
        const int tot = 64 * rows;
        for (int e0 = tid; e0 < tot; e0 += 64 * R64_WAVES * R64_UNROLL) {
            double v[R64_UNROLL];
            int at[R64_UNROLL];
#pragma unroll
            for (int u = 0; u < R64_UNROLL; u++) {
                const int e = e0 + u * 64 * R64_WAVES;
                const bool ok = e < tot;
                const int p = ok ? e / rows : 0, r = ok ? e - p * rows : 0;
                const int rs = rslot[r], x = rs >> 16, i = rs & 0xFFFF;
                const int t = tm[x * 64 + p];
                const int a = sa[x];
                const GLOBAL_AS double *col = as_global(A.tab) + as_global(A.tab_off)[a] + (int64_t)(t < 0 ? 0 : t) * sno[x];
                const double val = col[i];
                v[u] = t >= 0 ? val : 0.0;
                at[u] = ok ? r * R64_LD + p : -1;
            }
#pragma unroll
            for (int u = 0; u < R64_UNROLL; u++)
                if (at[u] >= 0) phi[at[u]] = v[u];
        }
        __syncthreads();
        z = rho64_pairs<true>(A, B, lds, L, lane, wave, k);
    } else {
        z = rho64_pairs<false>(A, B, lds, L, lane, wave, k);
    }
    // 5. the waves' sums in wave order
    part[wave * 64 + lane] = z;
    __syncthreads();
    if (wave == 0) {
        double r = part[lane];
#pragma unroll
        for (int w = 1; w < R64_WAVES; w++) r += part[w * 64 + lane];
        const int x1 = B.g1 + p1, x2 = B.g2 + p2, x3 = B.g3 + p3;
        if (x1 < A.n1 && x2 < A.n2 && x3 < A.n3) as_global(A.rho)[((int64_t)x1 * A.n2 + x2) * A.n3 + x3] = r;
    }
}

// Table form of the FP64 block density: as k_rblk64 without the staged
// columns -- each wave reads the lanes' columns of its pairs' slots straight
// from the tables (L2 serves the repeats) and sixteen waves share one
// sub-block's pair list (wave w takes pairs w, w + W, ...), so the LDS holds
// only the point map, the masks and the pair list and the workgroup starts
// without a staging phase.  GaAs216: 10.4 ms against 12.2 ms staged and
// 12.2 ms for the best pair-list kernel.  PPL points per lane (a region of
// 4 x 4 x 4 PPL points) would feed PPL FMAs per W element read; two points
// took 213 VGPRs and ran at 13.1 ms with eight waves (27 ms with sixteen,
// spilling), so one point per lane is what is launched.  Also measured: one
// workgroup per 8^3 block whose eight sub-block waves walk the block's pair
// list in the same order, to share each W block in the scalar cache: 15.2
// ms (16.5 ms with a barrier every four pairs).
template <int NA, int NB, int PPL>
__device__ __forceinline__ void quad64p(const GLOBAL_AS double *__restrict__ M, const double (&a)[PPL][NA],
                                        const double (&b)[PPL][NB], double (&z)[PPL]) {
#pragma unroll
    for (int i = 0; i < NA; i++) {
        double y[PPL];
#pragma unroll
        for (int u = 0; u < PPL; u++) y[u] = 0.0;
#pragma unroll
        for (int j = 0; j < NB; j++) {
            const double m = M[i * NB + j];
#pragma unroll
            for (int u = 0; u < PPL; u++) y[u] = fma(m, b[u][j], y[u]);
        }
#pragma unroll
        for (int u = 0; u < PPL; u++) z[u] = fma(a[u][i], y[u], z[u]);
    }
}

template <int N, int PPL>
__device__ __forceinline__ void cols64p(const GLOBAL_AS double *__restrict__ tab, const int (&t)[PPL],
                                        double (&f)[PPL][N]) {
#pragma unroll
    for (int u = 0; u < PPL; u++) {
        const GLOBAL_AS double *c = tab + (int64_t)(t[u] < 0 ? 0 : t[u]) * N;
#pragma unroll
        for (int i = 0; i < N; i++) {
            const double v = c[i];
            f[u][i] = t[u] >= 0 ? v : 0.0;
        }
    }
}

template <int NX, int NY, int PPL>
__device__ __forceinline__ void pair64t(const GLOBAL_AS double *__restrict__ M, bool tr, bool self,
                                        const GLOBAL_AS double *tx_tab, const int (&tx)[PPL],
                                        const GLOBAL_AS double *ty_tab, const int (&ty)[PPL], double (&z)[PPL]) {
    double fx[PPL][NX];
    cols64p<NX, PPL>(tx_tab, tx, fx);
    if constexpr (NX == NY) {
        if (self) { quad64p<NX, NX, PPL>(M, fx, fx, z); return; }
    }
    double fy[PPL][NY];
    cols64p<NY, PPL>(ty_tab, ty, fy);
    if (tr) quad64p<NY, NX, PPL>(M, fy, fx, z);
    else quad64p<NX, NY, PPL>(M, fx, fy, z);
}

__device__ __forceinline__ double quad64t_rt(const GLOBAL_AS double *__restrict__ M, const GLOBAL_AS double *ta_tab,
                                             int ta, int na, const GLOBAL_AS double *tb_tab, int tb, int nb) {
    if (ta < 0 || tb < 0) return 0.0;
    const GLOBAL_AS double *ca = ta_tab + (int64_t)ta * na, *cb = tb_tab + (int64_t)tb * nb;
    double z = 0.0;
    for (int i = 0; i < na; i++) {
        double y = 0.0;
        for (int j = 0; j < nb; j++) y = fma(M[i * nb + j], cb[j], y);
        z = fma(ca[i], y, z);
    }
    return z;
}

template <int PPL, int W>
__global__ __launch_bounds__(64 * W) void k_rblk64t(Rho64Args A) {
    extern __shared__ __align__(16) unsigned char lds[];
    constexpr int NREG = 8 / PPL;  // regions per block
    const int64_t nreg = A.nsub / 8 * NREG;
    const int64_t per = (nreg + 7) / 8;
    const int64_t j = blockIdx.x / 8, s = (int64_t)(blockIdx.x % 8) * per + j;
    if (j >= per || s >= nreg) return;
    const BlockDev B = as_global(A.blocks)[s / NREG];
    const int rg = (int)(s % NREG);
    const int ns = B.nslot, ms = A.max_slots;
    int32_t *tm = reinterpret_cast<int32_t *>(lds);                       // [slot][PPL][64]
    uint64_t *mask = reinterpret_cast<uint64_t *>(tm + ms * 64 * PPL);    // [slot][PPL]
    int32_t *pairs = reinterpret_cast<int32_t *>(mask + ms * PPL);
    double *part = reinterpret_cast<double *>(lds + ((ms * 64 * PPL * 4 + ms * PPL * 8 + ms * (ms + 1) / 2 * 4 + 7) / 8 * 8));
    int32_t *ctl = reinterpret_cast<int32_t *>(part + W * 64 * PPL);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    // the lane's points: region rg is sub-cube rg (PPL 1) or the 4 x 4 x 8
    // column (rg >> 1, rg & 1) (PPL 2)
    const int s1 = PPL == 1 ? rg >> 2 : rg >> 1, s2 = PPL == 1 ? (rg >> 1) & 1 : rg & 1;
    const int p1 = 4 * s1 + (lane >> 4), p2 = 4 * s2 + ((lane >> 2) & 3);
    int kk[PPL], p3[PPL];
#pragma unroll
    for (int u = 0; u < PPL; u++) {
        p3[u] = 4 * (PPL == 1 ? (rg & 1) : u) + (lane & 3);
        kk[u] = (p1 * BLK + p2) * BLK + p3[u];
    }
    const GLOBAL_AS int32_t *map = as_global(A.sph_map) + (int64_t)B.slot_off * BLK_PTS;
    for (int x = wave; x < ns; x += W) {
#pragma unroll
        for (int u = 0; u < PPL; u++) {
            const int t = map[(int64_t)x * BLK_PTS + kk[u]];
            tm[(x * PPL + u) * 64 + lane] = t;
            const uint64_t m = __ballot(t >= 0);
            if (lane == 0) mask[x * PPL + u] = m;
        }
    }
    __syncthreads();
    if (wave == 0) {
        int np = 0;
        for (int x = 0; x < ns; x++) {
            uint64_t any = 0;
#pragma unroll
            for (int u = 0; u < PPL; u++) any |= mask[x * PPL + u];
            if (!any) continue;
            for (int y0 = x; y0 < ns; y0 += 64) {
                const int y = y0 + lane;
                bool act = false;
                if (y < ns) {
                    uint64_t both = 0;
#pragma unroll
                    for (int u = 0; u < PPL; u++) both |= mask[y * PPL + u] & mask[x * PPL + u];
                    act = both != 0 && as_global(A.spair)[B.sp_off + (int64_t)x * ns + y].hoff >= 0;
                }
                const uint64_t bal = __ballot(act);
                if (act) pairs[np + __popcll(bal & ((1ull << lane) - 1))] = x << 16 | y;
                np += __popcll(bal);
            }
        }
        if (lane == 0) ctl[0] = np;
    }
    __syncthreads();
    const int np = ctl[0];
    const GLOBAL_AS double *Deff = as_global(A.Deff);
    const GLOBAL_AS SlotPair *sp = as_global(A.spair) + B.sp_off;
    const GLOBAL_AS int16_t *sno = as_global(A.slotno) + B.slot_off;
    const GLOBAL_AS int32_t *sa = as_global(A.sph_atom) + B.slot_off;
    double z[PPL];
#pragma unroll
    for (int u = 0; u < PPL; u++) z[u] = 0.0;
    for (int q = wave; q < np; q += W) {
        const int xy = __builtin_amdgcn_readfirstlane(pairs[q]);
        const int x = xy >> 16, y = xy & 0xFFFF;
        const SlotPair P = sp[x * ns + y];
        const GLOBAL_AS double *M = Deff + __builtin_amdgcn_readfirstlane(P.hoff);
        const bool tr = (__builtin_amdgcn_readfirstlane(P.meta) >> 6) != 0, self = x == y;
        const int nx = __builtin_amdgcn_readfirstlane(sno[x]), ny = __builtin_amdgcn_readfirstlane(sno[y]);
        const GLOBAL_AS double *xt = as_global(A.tab) + as_global(A.tab_off)[__builtin_amdgcn_readfirstlane(sa[x])];
        const GLOBAL_AS double *yt = as_global(A.tab) + as_global(A.tab_off)[__builtin_amdgcn_readfirstlane(sa[y])];
        int tx[PPL], ty[PPL];
#pragma unroll
        for (int u = 0; u < PPL; u++) {
            tx[u] = tm[(x * PPL + u) * 64 + lane];
            ty[u] = tm[(y * PPL + u) * 64 + lane];
        }
#define R64T_CASE(NX, NY)                                              \
    if (nx == NX && ny == NY) {                                        \
        pair64t<NX, NY, PPL>(M, tr, self, xt, tx, yt, ty, z);          \
        continue;                                                      \
    }
        NO_CASES(R64T_CASE)
#undef R64T_CASE
#pragma unroll
        for (int u = 0; u < PPL; u++)
            z[u] += tr ? quad64t_rt(M, yt, ty[u], ny, xt, tx[u], nx) : quad64t_rt(M, xt, tx[u], nx, yt, ty[u], ny);
    }
#pragma unroll
    for (int u = 0; u < PPL; u++) part[(wave * PPL + u) * 64 + lane] = z[u];
    __syncthreads();
    if (wave < PPL) {
        const int u = wave;
        double r = part[u * 64 + lane];
#pragma unroll
        for (int w = 1; w < W; w++) r += part[(w * PPL + u) * 64 + lane];
        const int x1 = B.g1 + p1, x2 = B.g2 + p2, x3 = B.g3 + p3[u];
        if (x1 < A.n1 && x2 < A.n2 && x3 < A.n3) as_global(A.rho)[((int64_t)x1 * A.n2 + x2) * A.n3 + x3] = r;
    }
}
template <int PPL, int W>
size_t r64t_lds(int ms) {
    return (size_t)((ms * 64 * PPL * 4 + ms * PPL * 8 + ms * (ms + 1) / 2 * 4 + 7) / 8 * 8) + W * 64 * PPL * 8 + 16;
}

// Matrix-core form of the FP64 block density.  Per 4x4x4 sub-block: the point
// map and presence masks in LDS and the present slots' columns staged in LDS
// ([row][point], stride R64_LD).  The slot pairs (x <= y) that share a point
// of the sub-block are listed y-major, and for each slot y the W blocks of its
// pairs are stacked row-wise (row (x, i): orbital i of x, for every listed x
// in slot order), so that
//   Q_y = W[stacked rows][orbitals of y] . Phi_y[orbitals of y][64 points]
// is one product on v_mfma_f64_16x16x4f64 whose 16-row tiles hold the rows
// of several pairs (each pair alone would pad its 13-19 rows to 32), and
//   rho(p) += sum over the stacked rows (x, i) of phi_{x,i}(p) Q_y[(x, i)][p]
// on the vector units.  The lists are built by all waves (one slot y per wave
// and pass, lane scans for the offsets); the work items are the 16-row tiles
// (y, tile), dealt round robin to the waves, each wave loading the W operands
// of its next tile while it multiplies the current one.  A lane holds the
// output rows kq, kq + 4, kq + 8, kq + 12 of a tile (kq = lane / 16) for the
// points lane % 16 + 16 nt and sums them over its tiles; the four lanes of a
// point and then the waves are added in a fixed order (deterministic).  A
// sub-block whose rows or stacked rows exceed the launch's LDS takes a vector
// path (one point per lane, exact FMAs, columns read from the tables).
typedef double d4_64 __attribute__((ext_vector_type(4)));

struct R64mLds {
    int phi, tm, mask, amask, tbase, sno, roff, yinfo, pairs, psp, rslot, items, stk, ctl, total;
    int item_cap, stack_cap, npair_cap;
};
__host__ __device__ inline R64mLds r64m_lds(int rows_cap, int ms, int stack_cap) {
    R64mLds L;
    L.stack_cap = stack_cap;
    L.item_cap = stack_cap / 16 + ms;
    L.npair_cap = ms * (ms + 1) / 2;
    const int phi = rows_cap * R64_LD * 8, part = R64_WAVES * 64 * 8;  // the wave sums reuse the column space
    L.phi = 0;
    L.tm = (phi > part ? phi : part);
    L.mask = L.tm + ms * 64 * 4;
    L.mask = (L.mask + 7) / 8 * 8;
    L.amask = L.mask + ms * 8;
    L.tbase = L.amask + ms * 8;
    L.sno = L.tbase + ms * 8;
    L.roff = L.sno + ms * 4;
    L.yinfo = L.roff + ms * 4;                 // per slot y: stacked rows, first stacked row, first item, first pair
    L.pairs = L.yinfo + ms * 4 * 4;            // x << 16 | y, y-major
    L.psp = L.pairs + L.npair_cap * 4;         // the slot pair of each listed pair
    L.psp = (L.psp + 7) / 8 * 8;
    L.rslot = L.psp + L.npair_cap * 8;         // staged row: x << 16 | i
    L.items = L.rslot + rows_cap * 4;          // y << 8 | tile
    L.stk = L.items + L.item_cap * 4;          // stacked row: listed pair << 15 | i << 10 | staged row
    L.ctl = L.stk + stack_cap * 4;
    L.ctl = (L.ctl + 15) / 16 * 16;
    L.total = L.ctl + 16;
    return L;
}

// exclusive prefix sum over the 64 lanes of a wave (and the total)
__device__ __forceinline__ int wave_scan_excl(int v, int lane, int &total) {
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    total = __shfl(x, 63, 64);
    return x - v;
}

// W operands of one 16-row tile: lane (kq, r16) holds W[row 16 mt + r16][4 ks + kq],
// and the staged rows of its output rows 16 mt + kq + 4 v (-1: past the end)
struct Tile64 {
    int y, s0, nr, ny, mt;
};
__device__ __forceinline__ Tile64 tile64(const int32_t *items, const int32_t *yinfo, const int32_t *sno, int q) {
    Tile64 T;
    const int it = __builtin_amdgcn_readfirstlane(items[q]);
    T.y = it >> 8;
    T.mt = it & 255;
    T.nr = __builtin_amdgcn_readfirstlane(yinfo[4 * T.y]);
    T.s0 = __builtin_amdgcn_readfirstlane(yinfo[4 * T.y + 1]);
    T.ny = __builtin_amdgcn_readfirstlane(sno[T.y]);
    return T;
}
__device__ __forceinline__ void tile64_w(const Tile64 &T, const int32_t *stk, const SlotPair *psp,
                                         const GLOBAL_AS double *__restrict__ Deff, int kq, int r16, double (&a)[8],
                                         int (&sr)[4]) {
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int rr = 16 * T.mt + kq + 4 * v;
        sr[v] = rr < T.nr ? (stk[T.s0 + rr] & 1023) : -1;
    }
    const int r = 16 * T.mt + r16;
    const bool rv = r < T.nr;
    const int ev = rv ? stk[T.s0 + r] : 0;
    const SlotPair P = psp[ev >> 15];
    const int ia = (ev >> 10) & 31, ld = P.meta & 63;
    const bool tr = (P.meta >> 6) != 0;
    const GLOBAL_AS double *M = Deff + (rv ? P.hoff : 0);
    const int kt = (T.ny + 3) >> 2;
#pragma unroll
    for (int ks = 0; ks < 8; ks++) {
        const int j = 4 * ks + kq;
        a[ks] = 0.0;
        if (ks < kt && rv && j < T.ny) a[ks] = M[tr ? j * ld + ia : ia * ld + j];
    }
}

__global__ __launch_bounds__(64 * R64_WAVES) void k_rblk64m(Rho64Args A, int stack_cap) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int64_t per = (A.nsub + 7) / 8;
    const int64_t jb = blockIdx.x / 8, s = (int64_t)(blockIdx.x % 8) * per + jb;
    if (jb >= per || s >= A.nsub) return;
    const BlockDev B = as_global(A.blocks)[s / 8];
    const int sc = (int)(s % 8);
    const int ns = B.nslot;  // <= 45 (MAX_SLOT_PAIRS): one lane per slot
    const R64mLds L = r64m_lds(A.rows_cap, A.max_slots, stack_cap);
    double *phi = reinterpret_cast<double *>(lds + L.phi);
    double *part = reinterpret_cast<double *>(lds + L.phi);
    int32_t *tm = reinterpret_cast<int32_t *>(lds + L.tm);
    uint64_t *mask = reinterpret_cast<uint64_t *>(lds + L.mask);
    uint64_t *amask = reinterpret_cast<uint64_t *>(lds + L.amask);
    int64_t *tbase = reinterpret_cast<int64_t *>(lds + L.tbase);
    int32_t *sno = reinterpret_cast<int32_t *>(lds + L.sno);
    int32_t *roff = reinterpret_cast<int32_t *>(lds + L.roff);
    int32_t *yinfo = reinterpret_cast<int32_t *>(lds + L.yinfo);
    int32_t *pairs = reinterpret_cast<int32_t *>(lds + L.pairs);
    SlotPair *psp = reinterpret_cast<SlotPair *>(lds + L.psp);
    int32_t *rslot = reinterpret_cast<int32_t *>(lds + L.rslot);
    int32_t *items = reinterpret_cast<int32_t *>(lds + L.items);
    int32_t *stk = reinterpret_cast<int32_t *>(lds + L.stk);
    int32_t *ctl = reinterpret_cast<int32_t *>(lds + L.ctl);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int p1 = 4 * (sc >> 2) + (lane >> 4), p2 = 4 * ((sc >> 1) & 1) + ((lane >> 2) & 3), p3 = 4 * (sc & 1) + (lane & 3);
    const int k = (p1 * BLK + p2) * BLK + p3;
    const GLOBAL_AS SlotPair *sp = as_global(A.spair) + B.sp_off;
    // 1. table columns (point index = lane), presence masks, orbital counts, table bases
    const GLOBAL_AS int32_t *map = as_global(A.sph_map) + (int64_t)B.slot_off * BLK_PTS;
    for (int x = wave; x < ns; x += R64_WAVES) {
        const int t = map[(int64_t)x * BLK_PTS + k];
        tm[x * 64 + lane] = t;
        const uint64_t m = __ballot(t >= 0);
        if (lane == 0) mask[x] = m;
    }
    if (wave == R64_WAVES - 1 && lane < ns) {
        sno[lane] = as_global(A.slotno)[B.slot_off + lane];
        tbase[lane] = as_global(A.tab_off)[as_global(A.sph_atom)[B.slot_off + lane]];
    }
    __syncthreads();
    // 2. per slot y (one wave each, lane = x): the slots x <= y of its listed
    // pairs; the slot pairs are read once here (unconditionally, clamped x)
    // and kept for the lists of step 4, which walks the same (wave, y) passes
    static_assert(MAX_SLOT_PAIRS < 46 * 46, "a block holds at most 45 sphere images");
    constexpr int YPASS = (45 + R64_WAVES - 1) / R64_WAVES;
    SlotPair spk[YPASS];
#pragma unroll
    for (int jp = 0; jp < YPASS; jp++) {
        const int y = wave + jp * R64_WAVES;
        if (y >= ns) break;
        const uint64_t my = mask[y];
        const int x = lane;
        const int xc = x <= y ? x : y;
        const SlotPair P = sp[(int64_t)xc * ns + y];
        spk[jp] = P;
        const bool act = my != 0 && x <= y && (mask[xc] & my) != 0 && P.hoff >= 0;
        const uint64_t bal = __ballot(act);
        int nr;
        wave_scan_excl(act ? sno[x] : 0, lane, nr);
        if (lane == 0) { amask[y] = bal; yinfo[4 * y] = nr; }
    }
    __syncthreads();
    // 3. offsets (one wave, lane = slot): staged rows, stacked rows, items, pairs
    if (wave == 0) {
        const int y = lane;
        const bool in = y < ns;
        int rows, nst, nit, npr;
        const int ro = wave_scan_excl(in && mask[in ? y : 0] ? sno[in ? y : 0] : 0, lane, rows);
        const int nr = in ? yinfo[4 * y] : 0;
        const int st0 = wave_scan_excl(nr, lane, nst);
        const int it0 = wave_scan_excl((nr + 15) >> 4, lane, nit);
        const int pp0 = wave_scan_excl(in ? __popcll(amask[y]) : 0, lane, npr);
        if (in) {
            roff[y] = ro;
            yinfo[4 * y + 1] = st0;
            yinfo[4 * y + 2] = it0;
            yinfo[4 * y + 3] = pp0;
        }
        if (lane == 0) {
            ctl[0] = npr;
            ctl[1] = rows;
            ctl[2] = nit;
            ctl[3] = (rows <= A.rows_cap && nst <= stack_cap && nit <= L.item_cap) ? 1 : 0;
        }
    }
    __syncthreads();
    const int rows = ctl[1];
    const bool mm = ctl[3] != 0;
    // 4. the lists (one slot y per wave and pass) and the staged rows
#pragma unroll
    for (int jp = 0; jp < YPASS; jp++) {
        const int y = wave + jp * R64_WAVES;
        if (y >= ns) break;
        const uint64_t am = amask[y];
        if (!am) continue;
        const int nr = yinfo[4 * y], st0 = yinfo[4 * y + 1], it0 = yinfo[4 * y + 2], pp0 = yinfo[4 * y + 3];
        const int x = lane;
        const bool act = x < ns && ((am >> x) & 1) != 0;
        int tot;
        const int cum = wave_scan_excl(act ? sno[x] : 0, lane, tot);
        const int q = pp0 + __popcll(am & ((1ull << lane) - 1));
        if (act) {
            pairs[q] = x << 16 | y;
            psp[q] = spk[jp];
            if (mm) {
                const int nx = sno[x], r0 = roff[x];
                for (int i = 0; i < nx; i++) stk[st0 + cum + i] = q << 15 | i << 10 | (r0 + i);
            }
        }
        if (mm)
            for (int t = lane; t < ((nr + 15) >> 4); t += 64) items[it0 + t] = y << 8 | t;
    }
    if (mm && tid < ns && mask[tid]) {
        const int nx = sno[tid], r0 = roff[tid];
        for (int i = 0; i < nx; i++) rslot[r0 + i] = tid << 16 | i;
    }
    __syncthreads();
    const GLOBAL_AS double *Deff = as_global(A.Deff);
    double z = 0.0;
    if (mm) {
        // 5. stage the present slots' columns: thread = (point, row lane), so
        // 16 lanes read 16 orbitals of one point (contiguous in the table)
        {
            const int p = tid >> 4, rl = tid & 15;
            const GLOBAL_AS double *tab = as_global(A.tab);
            for (int r0 = 0; r0 < rows; r0 += 16 * R64_UNROLL) {
                // unconditional loads from clamped rows and columns, zeroed
                // after, so that they issue together
                double v[R64_UNROLL];
#pragma unroll
                for (int u = 0; u < R64_UNROLL; u++) {
                    const int r = r0 + 16 * u + rl;
                    const int rs = rslot[r < rows ? r : rows - 1], x = rs >> 16, i = rs & 0xFFFF;
                    const int t = tm[x * 64 + p];
                    const double val = tab[tbase[x] + (int64_t)(t >= 0 ? t : 0) * sno[x] + i];
                    v[u] = t >= 0 ? val : 0.0;
                }
#pragma unroll
                for (int u = 0; u < R64_UNROLL; u++) {
                    const int r = r0 + 16 * u + rl;
                    if (r < rows) phi[r * R64_LD + p] = v[u];
                }
            }
        }
        __syncthreads();
        // 6. the tiles on the matrix cores
        const int nit = ctl[2];
        const int kq = lane >> 4, r16 = lane & 15;
        double zs[4] = {0.0, 0.0, 0.0, 0.0};
        double a[8];
        int sr[4];
        Tile64 T{0, 0, 0, 0, 0};
        if (wave < nit) {
            T = tile64(items, yinfo, sno, wave);
            tile64_w(T, stk, psp, Deff, kq, r16, a, sr);
        }
        for (int q = wave; q < nit; q += R64_WAVES) {
            const Tile64 C = T;
            double ac[8];
            int cr[4];
#pragma unroll
            for (int ks = 0; ks < 8; ks++) ac[ks] = a[ks];
#pragma unroll
            for (int v = 0; v < 4; v++) cr[v] = sr[v];
            // the next tile's W operands load under this tile's products
            if (q + R64_WAVES < nit) {
                T = tile64(items, yinfo, sno, q + R64_WAVES);
                tile64_w(T, stk, psp, Deff, kq, r16, a, sr);
            }
            const int kt = __builtin_amdgcn_readfirstlane((C.ny + 3) >> 2);
            d4_64 acc[4];
#pragma unroll
            for (int nt = 0; nt < 4; nt++) acc[nt] = d4_64{0.0, 0.0, 0.0, 0.0};
            const double *py = phi + (int64_t)__builtin_amdgcn_readfirstlane(roff[C.y]) * R64_LD + r16;
            // B operands: unconditional LDS reads (a clamped row, zero past
            // the slot's orbitals) so that they issue ahead of the products
#pragma unroll
            for (int ks = 0; ks < 8; ks++) {
                if (ks < kt) {
                    const int j = 4 * ks + kq;
                    const bool jv = j < C.ny;
                    const double *pj = py + (jv ? j : C.ny - 1) * R64_LD;
                    double b[4];
#pragma unroll
                    for (int nt = 0; nt < 4; nt++) {
                        const double bv = pj[16 * nt];
                        b[nt] = jv ? bv : 0.0;
                    }
#pragma unroll
                    for (int nt = 0; nt < 4; nt++)
                        acc[nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(ac[ks], b[nt], acc[nt], 0, 0, 0);
                }
            }
            // the lane's output rows 16 mt + kq + 4 v against phi_x at its
            // points (rows past the stack are zero in Q: any staged row serves)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const double *px = phi + (cr[v] >= 0 ? cr[v] : 0) * R64_LD + r16;
#pragma unroll
                for (int nt = 0; nt < 4; nt++) zs[nt] = fma(px[16 * nt], acc[nt][v], zs[nt]);
            }
        }
        // the four lanes of a point: rows kq = 0..3 (the same sums in every lane)
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            zs[nt] += __shfl_xor(zs[nt], 16, 64);
            zs[nt] += __shfl_xor(zs[nt], 32, 64);
        }
        z = kq == 0 ? zs[0] : kq == 1 ? zs[1] : kq == 2 ? zs[2] : zs[3];  // point 16 kq + r16 = lane
        __syncthreads();  // the columns are no longer read: their space takes the wave sums
    } else {
        // one point per lane, columns read from the tables, loops over the
        // orbitals (exact FMAs; only sub-blocks beyond the launch's LDS)
        const int np = ctl[0];
        const GLOBAL_AS double *tab = as_global(A.tab);
        for (int q = wave; q < np; q += R64_WAVES) {
            const int xy = __builtin_amdgcn_readfirstlane(pairs[q]);
            const int x = xy >> 16, y = xy & 0xFFFF;
            const SlotPair P = psp[q];
            const GLOBAL_AS double *M = Deff + __builtin_amdgcn_readfirstlane(P.hoff);
            const bool tr = (__builtin_amdgcn_readfirstlane(P.meta) >> 6) != 0;
            const int nx = __builtin_amdgcn_readfirstlane(sno[x]), ny = __builtin_amdgcn_readfirstlane(sno[y]);
            const GLOBAL_AS double *xt = tab + tbase[x];
            const GLOBAL_AS double *yt = tab + tbase[y];
            const int tx = tm[x * 64 + lane], ty = tm[y * 64 + lane];
            z += tr ? quad64t_rt(M, yt, ty, ny, xt, tx, nx) : quad64t_rt(M, xt, tx, nx, yt, ty, ny);
        }
    }
    // 7. the waves' sums in wave order
    part[wave * 64 + lane] = z;
    __syncthreads();
    if (wave == 0) {
        double r = part[lane];
#pragma unroll
        for (int w = 1; w < R64_WAVES; w++) r += part[w * 64 + lane];
        const int x1 = B.g1 + p1, x2 = B.g2 + p2, x3 = B.g3 + p3;
        if (x1 < A.n1 && x2 < A.n2 && x3 < A.n3) as_global(A.rho)[((int64_t)x1 * A.n2 + x2) * A.n3 + x3] = r;
    }
}

int split_density_f64_enqueue(lcpao_ctx *ctx, int nspin, const double *const *d_DM, int64_t dm_offset,
                              double *const *d_rho, int mode) {
    BlockPlan *P = ctx->plan;
    if (!P) { ctx->err = "the block plan is not built"; return LCPAO_ESTATE; }
    if (mode < 0 || mode > 2) { ctx->err = "unknown FP64 block density form"; return LCPAO_EINVAL; }
    const bool staged = mode == 0, mfma = mode == 2;
    // LDS: the staged columns of the largest sub-block, as far as 160 KB
    // allow (none when the columns are read from the tables); the
    // matrix-core form also holds the stacked rows (up to 2 bytes each)
    const int ms = std::max(1, P->max_slots);
    int cap = staged || mfma ? std::max(1, P->max_sub_rows) : 0;
    int scap = std::min(ms * (ms + 1) / 2 * 32, 16384);
    if (mfma) {
        while (scap > 2048 && r64m_lds(cap, ms, scap).total > 160 * 1024) scap -= 64;
        while (cap > 1 && r64m_lds(cap, ms, scap).total > 160 * 1024) cap--;
    } else {
        while (cap > 1 && r64_lds(cap, ms).total > 160 * 1024) cap--;
    }
    const void *kern = staged ? reinterpret_cast<const void *>(k_rblk64)
                       : mfma ? reinterpret_cast<const void *>(k_rblk64m)
                              : reinterpret_cast<const void *>(k_rblk64t<1, R64_WAVES>);
    const size_t lds = staged ? (size_t)r64_lds(cap, ms).total
                       : mfma ? (size_t)r64m_lds(cap, ms, scap).total
                              : r64t_lds<1, R64_WAVES>(ms);
    if (lds > 160 * 1024) { ctx->err = "the FP64 block density does not fit the LDS for this geometry"; return LCPAO_ERANGE; }
    SPLIT_TRY(ctx, hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const int64_t nsub = P->nblocks * 8, per = (nsub + 7) / 8;
    for (int s = 0; s < nspin; s++) {
        if (!d_DM[s] || !d_rho[s]) { ctx->err = "null spin pointer"; return LCPAO_EINVAL; }
        if (ctx->ncanon > 0) {
            KTimer kt(ctx, KID_DEFF);
            launch_deff_f64(ctx, d_DM[s] + dm_offset, 0);
        }
        if (nsub > 0) {
            KTimer kt(ctx, KID_RHO);
            const Rho64Args A{P->d_blocks, P->d_sph_atom, P->d_sph_map, P->d_slotno, P->d_spair, ctx->d_tab_off,
                              ctx->d_tab, ctx->d_deff, d_rho[s], nsub, ctx->ngrid[0], ctx->ngrid[1], ctx->ngrid[2],
                              cap, ms, 0};
            if (staged)
                hipLaunchKernelGGL(k_rblk64, dim3((unsigned)(8 * per)), dim3(64 * R64_WAVES), lds, ctx->stream, A);
            else if (mfma)
                hipLaunchKernelGGL(k_rblk64m, dim3((unsigned)(8 * per)), dim3(64 * R64_WAVES), lds, ctx->stream, A,
                                   scap);
            else
                hipLaunchKernelGGL((k_rblk64t<1, R64_WAVES>), dim3((unsigned)(8 * per)), dim3(64 * R64_WAVES), lds,
                                   ctx->stream, A);
        }
        SPLIT_TRY(ctx, hipGetLastError());
    }
    return LCPAO_OK;
}

// ----------------------------------------------------------------------------
// block Force3 (bf16x3): per block, T = D' Phi on the matrix cores with D' the
// ordered-pair density blocks (row orbital i of slot x, column orbital j of
// slot y: D(pair x -> y)[i][j]), then per row i and direction d
//   g[i][d] = sum_p V(p) dphi_d[i](p) T[i](p)
// on the vector units, and S_A[d] = sum of g over the rows of A's slots.
// ----------------------------------------------------------------------------

// gradient tiles: element (d, row, point k = 32 c + 16 pt + k16) of a block
// at ((((c * 2 + pt) * 3 + d) * (R / 16) + row / 16) * 64 + lane) * 4 + k16 % 4
// with lane = 16 (k16 / 4) + row % 16: lane l of a wave reads 16 contiguous
// bytes, orbital 16 g + l % 16 at the four points 4 (l / 16) .. + 3 of a
// 16-point group, which are the points of its matrix-core outputs
__host__ __device__ __forceinline__ int64_t gradfrag_index(int c, int pt, int d, int row, int k16, int R) {
    return ((((int64_t)(c * 2 + pt) * 3 + d) * (R / 16) + row / 16) * 64 + (k16 / 4) * 16 + row % 16) * 4 + k16 % 4;
}

__global__ void k_dtiles(int64_t nsph, const int32_t *__restrict__ sph_block, const int32_t *__restrict__ sph_atom,
                         const int32_t *__restrict__ sph_row0, const int32_t *__restrict__ sph_map,
                         const BlockDev *__restrict__ blocks, const int64_t *__restrict__ tab_off,
                         const int32_t *__restrict__ no, const double *__restrict__ dx, const double *__restrict__ dy,
                         const double *__restrict__ dz, float *__restrict__ dtile) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nsph * BLK_PTS) return;
    const int64_t s = gid / BLK_PTS;
    const int k = (int)(gid - s * BLK_PTS);
    const int t = sph_map[gid];
    if (t < 0) return;  // the tiles are zeroed first
    const int a = sph_atom[s];
    const int NO = no[a];
    const BlockDev B = blocks[sph_block[s]];
    const int c = k / CHUNK, pt = (k % CHUNK) / 16, k16 = k % 16;
    const int64_t col = tab_off[a] + (int64_t)t * NO;
    float *out = dtile + 3 * B.tile_off;
    for (int i = 0; i < NO; i++) {
        const int row = sph_row0[s] + i;
        out[gradfrag_index(c, pt, 0, row, k16, B.rows)] = (float)dx[col + i];
        out[gradfrag_index(c, pt, 1, row, k16, B.rows)] = (float)dy[col + i];
        out[gradfrag_index(c, pt, 2, row, k16, B.rows)] = (float)dz[col + i];
    }
}

struct ForceArgs {
    const BlockDev *blocks;
    const uint16_t *thi, *tlo;     // point fragments of Phi
    const float *dtile;            // gradient tiles
    const int16_t *rowinfo;
    const SlotPair *fpair;         // ordered slot pairs
    const int16_t *slotno;
    const float *D;                // the channel's density-matrix slice (FP32)
    const double *V;
    const int32_t *base;           // per item: its first (item, slot) output
    double *out;                   // per (item, slot): the three sums
    int32_t n1, n2, n3, pad;
};

// LDS of the Force3 workgroup: the chunk planes (as the density kernel),
// V at the block's points, per-wave row sums, per-row FP64 sums, row
// descriptors, slot orbital counts, then the ordered slot pairs
constexpr int FB_VB_OFFSET = MAX_ROWS * RB_BYTES_PER_ROW;
constexpr int FB_SUM_OFFSET = FB_VB_OFFSET + BLK_PTS * 4;
constexpr int FB_TOT_OFFSET = FB_SUM_OFFSET + FB_WAVES * ST * 3 * 4;
constexpr int FB_RINFO_OFFSET = FB_TOT_OFFSET + MAX_ROWS * 3 * 8;
constexpr int FB_MAX_SLOTS = 64;
constexpr int FB_NO_OFFSET = FB_RINFO_OFFSET + MAX_ROWS * 2;
constexpr int FB_FP_OFFSET = FB_NO_OFFSET + FB_MAX_SLOTS * 2;
static_assert(FB_TOT_OFFSET % 8 == 0 && FB_FP_OFFSET % 8 == 0, "Force3 LDS areas are aligned");

size_t force_kernel_lds(int max_slots) { return FB_FP_OFFSET + (size_t)max_slots * max_slots * sizeof(SlotPair); }

// L2 warm-up of a chunk's gradient tiles (3 R lines of 128 bytes) by the
// workgroup's threads, as touch_chunk_n
struct GradTouch {
    uint32_t v[(3 * MAX_ROWS + 64 * FB_WAVES - 1) / (64 * FB_WAVES)];
};
__device__ __forceinline__ GradTouch touch_grad(const GLOBAL_AS float *__restrict__ g, int64_t off, int R, int tid) {
    GradTouch t;
#pragma unroll
    for (int k = 0; k < (int)(sizeof(t.v) / sizeof(t.v[0])); k++) {
        const int x = tid + k * 64 * FB_WAVES < 3 * R ? tid + k * 64 * FB_WAVES : 0;
        t.v[k] = *reinterpret_cast<const GLOBAL_AS uint32_t *>(g + off + (int64_t)x * 32);
    }
    return t;
}
__device__ __forceinline__ void release_grad(const GradTouch &t) {
#pragma unroll
    for (int k = 0; k < (int)(sizeof(t.v) / sizeof(t.v[0])); k++) asm volatile("" ::"v"(t.v[k]));
}

// One 512-thread workgroup per item (a block and up to 8 units of its S x S
// super-tiles), one unit per wave: super-tiles (si, sj) and (si, sj + 1).
// The wave holds D'[si][sj] and D'[si][sj + 1] as B-operand fragments (lane
// l: D'[64 si + 16 ct + l % 16][64 sj + 32 ks + 8 (l / 16) + e]) and
// multiplies the point fragments of Phi (rows sj, sj + 1) as the A operand,
// so that output element v of lane l is T[orbital 64 si + 16 ct + l % 16]
// [point 16 pt + 4 (l / 16) + v], summed over both super-tiles in one
// accumulator.  Each lane accumulates g for its four orbitals over its
// points; the lanes of an orbital are added at the end, the waves of a
// super-tile row in wave order (FP64), and the rows of a slot in row order:
// deterministic.
template <int NF>
__global__ __launch_bounds__(64 * FB_WAVES) void k_fblk_bf(const HItem *__restrict__ items, int32_t nitems,
                                                          ForceArgs A) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    if ((int)blockIdx.x >= nitems) return;
    const HItem it = items[blockIdx.x];
    const BlockDev B = as_global(A.blocks)[it.block];
    const GLOBAL_AS uint16_t *__restrict__ thi = as_global(A.thi);
    const GLOBAL_AS uint16_t *__restrict__ tlo = as_global(A.tlo);
    const GLOBAL_AS float *__restrict__ dtile = as_global(A.dtile) + 3 * B.tile_off;
    const GLOBAL_AS float *__restrict__ Dm = as_global(A.D);
    const int R = B.rows, S = R / ST, U = (S + 1) / 2, G = R / 16, G32 = R / 32;
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63, l16 = lane & 15,
              q = lane >> 4;
    u16x8 *TH = reinterpret_cast<u16x8 *>(lds_raw);
    u16x8 *TL = TH + G * 64;
    float *Vb = reinterpret_cast<float *>(lds_raw + FB_VB_OFFSET);
    float *rsum = reinterpret_cast<float *>(lds_raw + FB_SUM_OFFSET);     // [wave][64 rows][3]
    double *rtot = reinterpret_cast<double *>(lds_raw + FB_TOT_OFFSET);   // [row][3]
    int16_t *s_rinfo = reinterpret_cast<int16_t *>(lds_raw + FB_RINFO_OFFSET);
    int16_t *s_no = reinterpret_cast<int16_t *>(lds_raw + FB_NO_OFFSET);
    SlotPair *s_fp = reinterpret_cast<SlotPair *>(lds_raw + FB_FP_OFFSET);
    constexpr int NT = 64 * FB_WAVES;
    for (int x = tid; x < R; x += NT) s_rinfo[x] = as_global(A.rowinfo)[B.row_off + x];
    for (int x = tid; x < B.nslot; x += NT) s_no[x] = as_global(A.slotno)[B.slot_off + x];
    for (int x = tid; x < B.nslot * B.nslot; x += NT) s_fp[x] = as_global(A.fpair)[B.sp_off + x];
    for (int k = tid; k < BLK_PTS; k += NT) {
        const int x1 = B.g1 + k / (BLK * BLK), x2 = B.g2 + (k / BLK) % BLK, x3 = B.g3 + k % BLK;
        Vb[k] = (x1 < A.n1 && x2 < A.n2 && x3 < A.n3) ? (float)as_global(A.V)[((int64_t)x1 * A.n2 + x2) * A.n3 + x3]
                                                      : 0.f;
    }
    // the first two chunks are fetched into L2 while D' is gathered
    const int64_t gstride = (int64_t)6 * G * 64 * 4;  // floats of one chunk's gradient tiles
    __syncthreads();
    const bool has = wave < it.nst;
    int si = 0, sj = 0;
    if (has) {
        si = (it.st0 + wave) / U;
        sj = 2 * ((it.st0 + wave) % U);
    }
    const bool has2 = has && sj + 1 < S;  // the row's last unit of an odd S holds one super-tile
    // D' gathered straight into the B-operand fragments, 32 loads per batch
    bf16x8 wh[2][4][2], wl[2][4][2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        constexpr int GB = 4;
        const float keep = (h == 0 ? has : has2) ? 1.f : 0.f;
        const int cj0 = (sj + h < S ? sj + h : sj) * ST;
#pragma unroll
        for (int b2 = 0; b2 < 8; b2 += GB) {
            float dv[GB][8];
            uint32_t okbits = 0;
#pragma unroll
            for (int r2 = 0; r2 < GB; r2++) {
                const int ct = (b2 + r2) >> 1, ks = (b2 + r2) & 1;
                const int rinf = s_rinfo[si * ST + 16 * ct + l16];
                const int ra = rinf >> 5, ri = rinf & 31;
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const int cinf = s_rinfo[cj0 + 32 * ks + 8 * q + e];
                    const int cb = cinf >> 5, cj = cinf & 31;
                    const bool valid = rinf >= 0 && cinf >= 0;
                    const SlotPair P = s_fp[valid ? ra * B.nslot + cb : 0];
                    const bool ok = valid && P.hoff >= 0;
                    dv[r2][e] = Dm[ok ? P.hoff + ri * P.meta + cj : 0];  // unconditional: loads issue back to back
                    okbits |= (ok ? 1u : 0u) << (r2 * 8 + e);
                }
            }
#pragma unroll
            for (int r2 = 0; r2 < GB; r2++) {
                const int ct = (b2 + r2) >> 1, ks = (b2 + r2) & 1;
                u16x8 wh8, wl8;
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const float w = dv[r2][e] * (((okbits >> (r2 * 8 + e)) & 1u) ? keep : 0.f);
                    const __bf16 hb = (__bf16)w;
                    wh8[e] = __builtin_bit_cast(uint16_t, hb);
                    wl8[e] = __builtin_bit_cast(uint16_t, (__bf16)(w - (float)hb));
                }
                wh[h][ct][ks] = __builtin_bit_cast(bf16x8, wh8);
                wl[h][ct][ks] = __builtin_bit_cast(bf16x8, wl8);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    TouchN<NT> t0 = touch_chunk_n<NT>(thi, tlo, B.tile_off, R, tid);
    TouchN<NT> t1 = touch_chunk_n<NT>(thi, tlo, B.tile_off + (int64_t)G * FRAG, R, tid);
    GradTouch u0 = touch_grad(dtile, 0, R, tid);
    GradTouch u1 = touch_grad(dtile, gstride, R, tid);
    float g[4][3];
#pragma unroll
    for (int ct = 0; ct < 4; ct++)
#pragma unroll
        for (int d = 0; d < 3; d++) g[ct][d] = 0.f;
    u16x8 fh[NF], fl[NF];
    auto frag_of = [&](int f) {
        const int gg = wave + f * FB_WAVES;
        return gg < G ? gg : min(wave, G - 1);
    };
    auto prefetch = [&](int c) {
        const GLOBAL_AS uint16_t *sh = thi + B.tile_off + (int64_t)c * G * FRAG;
        const GLOBAL_AS uint16_t *sl = tlo + B.tile_off + (int64_t)c * G * FRAG;
#pragma unroll
        for (int f = 0; f < NF; f++) {  // unconditional loads: see touch_chunk
            const int gg = frag_of(f);
            fh[f] = *reinterpret_cast<const GLOBAL_AS u16x8 *>(sh + (gg * 64 + lane) * 8);
            fl[f] = *reinterpret_cast<const GLOBAL_AS u16x8 *>(sl + (gg * 64 + lane) * 8);
        }
    };
    const int sj2 = sj + 1 < S ? sj + 1 : sj;  // in bounds; its D' is zero when absent
    auto step = [&](int c, TouchN<NT> &slot, GradTouch &gslot) {
        __syncthreads();
#pragma unroll
        for (int f = 0; f < NF; f++) {
            const int gg = frag_of(f);
            TH[gg * 64 + lane] = fh[f];
            TL[gg * 64 + lane] = fl[f];
        }
        release_n(slot);
        release_grad(gslot);
        __syncthreads();
        prefetch(min(c + 1, NCHUNK - 1));
        const int c2 = min(c + 2, NCHUNK - 1);
        slot = touch_chunk_n<NT>(thi, tlo, B.tile_off + (int64_t)c2 * G * FRAG, R, tid);
        gslot = touch_grad(dtile, c2 * gstride, R, tid);
        const GLOBAL_AS f32x4 *dg = reinterpret_cast<const GLOBAL_AS f32x4 *>(dtile + c * gstride);
#pragma unroll 1
        for (int pt = 0; pt < 2; pt++) {
            bf16x8 bh[2][2], bl[2][2];
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int ks = 0; ks < 2; ks++) {
                    const int fi = pt * G32 + (h == 0 ? sj : sj2) * 2 + ks;
                    bh[h][ks] = __builtin_bit_cast(bf16x8, TH[fi * 64 + lane]);
                    bl[h][ks] = __builtin_bit_cast(bf16x8, TL[fi * 64 + lane]);
                }
            const f32x4 vv = *reinterpret_cast<const f32x4 *>(&Vb[c * CHUNK + pt * 16 + 4 * q]);
#pragma unroll
            for (int ct = 0; ct < 4; ct++) {
                __builtin_amdgcn_sched_barrier(0);
                f32x4 dd[3];
#pragma unroll
                for (int d = 0; d < 3; d++) dd[d] = dg[((pt * 3 + d) * G + si * 4 + ct) * 64 + lane];
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int h = 0; h < 2; h++)
#pragma unroll
                    for (int ks = 0; ks < 2; ks++) {
                        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[h][ks], wh[h][ct][ks], acc, 0, 0, 0);
                        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[h][ks], wl[h][ct][ks], acc, 0, 0, 0);
                        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[h][ks], wh[h][ct][ks], acc, 0, 0, 0);
                    }
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const float tv = acc[v] * vv[v];
#pragma unroll
                    for (int d = 0; d < 3; d++) g[ct][d] = fmaf(tv, dd[d][v], g[ct][d]);
                }
            }
        }
    };
    prefetch(0);
#pragma unroll
    for (int c = 0; c < NCHUNK; c += 2) {
        step(c, t0, u0);
        step(c + 1, t1, u1);
    }
    // sum over the four lanes of each orbital (lanes 16 apart) with the lane swaps
#pragma unroll
    for (int ct = 0; ct < 4; ct++)
#pragma unroll
        for (int d = 0; d < 3; d++) {
            float x = g[ct][d];
            const auto s32 = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
            x = __uint_as_float(s32[0]) + __uint_as_float(s32[1]);
            const auto s16 = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
            g[ct][d] = __uint_as_float(s16[0]) + __uint_as_float(s16[1]);
        }
    if (q == 0 && has)
#pragma unroll
        for (int ct = 0; ct < 4; ct++)
#pragma unroll
            for (int d = 0; d < 3; d++) rsum[(wave * ST + 16 * ct + l16) * 3 + d] = g[ct][d];
    __syncthreads();
    for (int x = tid; x < R * 3; x += NT) {
        const int r = x / 3, d = x % 3;
        double t = 0.0;
        for (int w = 0; w < it.nst; w++)
            if ((it.st0 + w) / U == r / ST) t += (double)rsum[(w * ST + r % ST) * 3 + d];
        rtot[x] = t;
    }
    __syncthreads();
    for (int x = tid; x < B.nslot * 3; x += NT) {
        const int xs = x / 3, d = x % 3;
        int r0 = 0;
        for (int y = 0; y < xs; y++) r0 += s_no[y];
        double t = 0.0;
        for (int i = 0; i < s_no[xs]; i++) t += rtot[(r0 + i) * 3 + d];
        A.out[(int64_t)(A.base[blockIdx.x] + xs) * 3 + d] = t;
    }
}

// per atom: the sums of its (item, slot) entries, x w, added into
// fsum[a][d]; one wave per atom, lane l adds entries l, l + 64, ... in item
// order, then a fixed butterfly (deterministic)
__global__ __launch_bounds__(256) void k_fblk_reduce(const int32_t *__restrict__ off, const int32_t *__restrict__ idx,
                                                     int natom, const double *__restrict__ out, double w,
                                                     double *__restrict__ fsum) {
    const int a = blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (a >= natom) return;  // whole waves
    double s[3] = {0.0, 0.0, 0.0};
    for (int u = off[a] + lane; u < off[a + 1]; u += 64)
        for (int d = 0; d < 3; d++) s[d] += out[3 * (int64_t)idx[u] + d];
    for (int o = 32; o >= 1; o >>= 1)
        for (int d = 0; d < 3; d++) s[d] += __shfl_xor(s[d], o, 64);
    if (lane == 0)
        for (int d = 0; d < 3; d++) fsum[3 * a + d] += w * s[d];
}

int split_force3_ready(lcpao_ctx *ctx) {
    int rc = split_tables(ctx, LCPAO_BF16X3);
    if (rc) return rc;
    if (ctx->plan->max_slots > FB_MAX_SLOTS) {
        ctx->err = "too many sphere images per block for the bf16x3 Force3 (" + std::to_string(ctx->plan->max_slots) +
                   ", at most " + std::to_string(FB_MAX_SLOTS) + ")";
        return LCPAO_ERANGE;
    }
    return LCPAO_OK;
}

int split_force3(lcpao_ctx *ctx, const double *d_V, const double *d_DM, double w, double *d_fsum) {
    int rc = split_force3_ready(ctx);
    if (rc) return rc;
    BlockPlan *P = ctx->plan;
    if (!P->dtile_ready) {
        const size_t n = (size_t)std::max<int64_t>(1, P->tile_elems) * 3;
        if (!P->d_dtile) SPLIT_TRY(ctx, hipMalloc(&P->d_dtile, n * sizeof(float)));
        if (!P->d_fout) SPLIT_TRY(ctx, hipMalloc(&P->d_fout, std::max<int64_t>(1, P->fout) * 3 * sizeof(double)));
        SPLIT_TRY(ctx, hipMemsetAsync(P->d_dtile, 0, n * sizeof(float), ctx->stream));
        const int64_t nthr = P->nsph * BLK_PTS;
        if (nthr > 0)
            hipLaunchKernelGGL(k_dtiles, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, ctx->stream, P->nsph,
                               P->d_sph_block, P->d_sph_atom, P->d_sph_row0, P->d_sph_map, P->d_blocks, ctx->d_tab_off,
                               ctx->d_no, ctx->d_dtab[0], ctx->d_dtab[1], ctx->d_dtab[2], P->d_dtile);
        SPLIT_TRY(ctx, hipGetLastError());
        SPLIT_TRY(ctx, hipStreamSynchronize(ctx->stream));
        P->dtile_ready = true;
    }
    launch_to_f32(ctx->stream, ctx->my_hsize, d_DM, ctx->d_deff32);
    if (!P->fitems.empty()) {
        KTimer kt(ctx, KID_F3);
        const ForceArgs A{P->d_blocks, P->d_thi, P->d_tlo, P->d_dtile, P->d_rowinfo, P->d_fpair, P->d_slotno,
                          ctx->d_deff32, d_V, P->d_fbase, P->d_fout, ctx->ngrid[0], ctx->ngrid[1], ctx->ngrid[2], 0};
        void (*kern)(const HItem *, int32_t, ForceArgs) = nullptr;
        switch (nfrag(P->max_rows)) {
            case 1: kern = k_fblk_bf<1>; break;
            case 2: kern = k_fblk_bf<2>; break;
            case 3: kern = k_fblk_bf<3>; break;
            case 4: kern = k_fblk_bf<4>; break;
            default: kern = k_fblk_bf<MAXF>; break;
        }
        const size_t lds = force_kernel_lds(P->max_slots);
        SPLIT_TRY(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        const unsigned ni = (unsigned)P->fitems.size();
        hipLaunchKernelGGL(kern, dim3(ni), dim3(64 * FB_WAVES), lds, ctx->stream, P->d_fitems, (int32_t)ni, A);
    }
    if (ctx->natom > 0) {
        KTimer kt(ctx, KID_F3RED);
        hipLaunchKernelGGL(k_fblk_reduce, dim3((ctx->natom + 3) / 4), dim3(256), 0, ctx->stream, P->d_fatom_off,
                           P->d_fatom_idx, ctx->natom, P->d_fout, w, d_fsum);
    }
    SPLIT_TRY(ctx, hipGetLastError());
    return LCPAO_OK;
}

}  // namespace lcpao

// ---------------------------------------------------------------------------
// Host-only layout check of the block forms (no device needed): the same
// bucketing of the grid-list entries into 8x8x8 blocks as the plan, and the
// same limits.
// ---------------------------------------------------------------------------
extern "C" int lcpao_grid_block_layout(const lcpao_geom *g, int32_t *max_rows, int32_t *max_slots, int64_t *nblocks) {
    using namespace lcpao;
    if (!g || g->natom <= 0 || !g->no || !g->gridn || !g->grid_list || !g->cell_list || g->ngrid[0] <= 0 ||
        g->ngrid[1] <= 0 || g->ngrid[2] <= 0)
        return LCPAO_EINVAL;
    const int N1 = g->ngrid[0], N2 = g->ngrid[1], N3 = g->ngrid[2];
    const int nb1 = (N1 + BLK - 1) / BLK, nb2 = (N2 + BLK - 1) / BLK, nb3 = (N3 + BLK - 1) / BLK;
    const int64_t NB = (int64_t)nb1 * nb2 * nb3;
    // (block, atom, cell) of every grid-list entry, unique per block
    std::vector<std::pair<int64_t, int64_t>> key;
    int maxno = 0;
    for (int a = 0; a < g->natom; a++) {
        maxno = std::max(maxno, (int)g->no[a]);
        if (g->gridn[a] > 0 && (!g->grid_list[a] || !g->cell_list[a])) return LCPAO_EINVAL;
        for (int t = 0; t < g->gridn[a]; t++) {
            const int gi = g->grid_list[a][t];
            if (gi < 0 || (int64_t)gi >= (int64_t)N1 * N2 * N3) return LCPAO_EINVAL;
            const int n1 = gi / (N2 * N3), n2 = (gi / N3) % N2, n3 = gi % N3;
            const int64_t b = ((int64_t)(n1 / BLK) * nb2 + (n2 / BLK)) * nb3 + (n3 / BLK);
            key.push_back({b, (int64_t)a * 1000000 + g->cell_list[a][t]});
        }
    }
    std::sort(key.begin(), key.end());
    key.erase(std::unique(key.begin(), key.end()), key.end());
    int mr = 0, ms = 0;
    for (size_t i = 0; i < key.size();) {
        size_t j = i;
        int r = 0, ns = 0;
        while (j < key.size() && key[j].first == key[i].first) { r += g->no[key[j].second / 1000000]; ns++; j++; }
        mr = std::max(mr, std::max(ST, (r + ST - 1) / ST * ST));
        ms = std::max(ms, ns);
        i = j;
    }
    if (max_rows) *max_rows = mr;
    if (max_slots) *max_slots = ms;
    if (nblocks) *nblocks = NB;
    const bool fits = maxno <= 32 && mr <= MAX_ROWS && ms * ms <= MAX_SLOT_PAIRS && ms <= FB_MAX_SLOTS;
    return fits ? LCPAO_OK : LCPAO_ERANGE;
}
