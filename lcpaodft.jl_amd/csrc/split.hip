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

#define SPLIT_TRY(ctx, call)                                                        \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                     \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);        \
            return LCPAO_EDEVICE;                                                   \
        }                                                                           \
    } while (0)

constexpr int MAX_ROWS = 768;    // R limit (LDS: 2*R*40*2 B)
constexpr int LROW = 40;         // LDS row stride of the bf16 planes (32 + 8 pad)
constexpr int LROW32 = 36;       // LDS row stride of fp32 tiles (32 + 4 pad)

__device__ __forceinline__ int st_index(int si, int sj, int S) { return si * S - (si * (si - 1)) / 2 + (sj - si); }

// position of point q (0..31) inside a fp32 chunk row: lane group q%4 of the
// 16x16x4 layout reads its 8 points contiguously
__host__ __device__ __forceinline__ int perm32(int q) { return (q & 3) * 8 + (q >> 2); }

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
            const int64_t idx = B.tile_off + ((int64_t)c * B.rows + row) * CHUNK + q;
            hi[idx] = __builtin_bit_cast(uint16_t, h);
            lo[idx] = __builtin_bit_cast(uint16_t, l);
            const int64_t tdx = B.tile_off + ((int64_t)c * CHUNK + q) * B.rows + row;
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
template <int MODE>
__device__ __forceinline__ void supertile(const uint16_t *__restrict__ Ahi, const uint16_t *__restrict__ Alo,
                                          const float *__restrict__ A32, const float *__restrict__ Vs, int si, int sj,
                                          int l16, int q, f32x4 *acc) {
    if constexpr (MODE == LCPAO_BF16X3) {
        bf16x8 ah[4], al[4];
#pragma unroll
        for (int rt = 0; rt < 4; rt++) {
            const int row = si * ST + rt * 16 + l16;
            ah[rt] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(&Ahi[row * LROW + q * 8]));
            al[rt] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(&Alo[row * LROW + q * 8]));
        }
#pragma unroll
        for (int ct = 0; ct < 4; ct++) {
            const int row = sj * ST + ct * 16 + l16;
            const bf16x8 xh = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(&Ahi[row * LROW + q * 8]));
            const bf16x8 xl = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(&Alo[row * LROW + q * 8]));
            bf16x8 bh, bl;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const float x = ((float)xh[e] + (float)xl[e]) * Vs[q * 8 + e];
                const __bf16 h = (__bf16)x;
                bh[e] = h;
                bl[e] = (__bf16)(x - (float)h);
            }
#pragma unroll
            for (int rt = 0; rt < 4; rt++) {
                f32x4 c = acc[rt * 4 + ct];
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[rt], bh, c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rt], bl, c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rt], bh, c, 0, 0, 0);
                acc[rt * 4 + ct] = c;
            }
        }
    } else {
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
                f32x4 c = acc[rt * 4 + ct];
#pragma unroll
                for (int s = 0; s < 8; s++) c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rt][s], b[s], c, 0, 0, 0);
                acc[rt * 4 + ct] = c;
            }
        }
    }
}

__device__ __forceinline__ void st_decode(int st, int S, int &si, int &sj) {
    si = 0;
    while (st >= S - si) { st -= S - si; si++; }
    sj = si + st;
}

template <int MODE>
__global__ __launch_bounds__(64 * HB_WAVES) void k_hblk(const HItem *__restrict__ items, const BlockDev *__restrict__ blocks,
                                                       const uint16_t *__restrict__ hi, const uint16_t *__restrict__ lo,
                                                       const float *__restrict__ t32, const double *__restrict__ V,
                                                       int n1, int n2, int n3, float *__restrict__ part) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const HItem it = items[blockIdx.x];
    const BlockDev B = blocks[it.block];
    const int R = B.rows, S = R / ST;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, l16 = lane & 15, q = lane >> 4;
    uint16_t *Ahi = reinterpret_cast<uint16_t *>(lds_raw);
    uint16_t *Alo = Ahi + R * LROW;
    float *A32 = reinterpret_cast<float *>(lds_raw);
    float *Vs = MODE == LCPAO_BF16X3 ? reinterpret_cast<float *>(Alo + R * LROW) : A32 + R * LROW32;
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
        if constexpr (MODE == LCPAO_BF16X3) {
            const uint16_t *sh = hi + B.tile_off + (int64_t)c * R * CHUNK;
            const uint16_t *sl = lo + B.tile_off + (int64_t)c * R * CHUNK;
            for (int e = tid * 8; e < R * CHUNK; e += 64 * HB_WAVES * 8) {
                const int row = e / CHUNK, col = e % CHUNK;
                *reinterpret_cast<u16x8 *>(&Ahi[row * LROW + col]) = *reinterpret_cast<const u16x8 *>(&sh[e]);
                *reinterpret_cast<u16x8 *>(&Alo[row * LROW + col]) = *reinterpret_cast<const u16x8 *>(&sl[e]);
            }
        } else {
            const float *s32 = t32 + B.tile_off + (int64_t)c * R * CHUNK;
            for (int e = tid * 4; e < R * CHUNK; e += 64 * HB_WAVES * 4) {
                const int row = e / CHUNK, col = e % CHUNK;
                *reinterpret_cast<f32x4 *>(&A32[row * LROW32 + col]) = *reinterpret_cast<const f32x4 *>(&s32[e]);
            }
        }
        if (tid < CHUNK) {
            const int k = c * CHUNK + tid;
            const int x1 = B.g1 + k / (BLK * BLK), x2 = B.g2 + (k / BLK) % BLK, x3 = B.g3 + k % BLK;
            const float v = (x1 < n1 && x2 < n2 && x3 < n3) ? (float)V[((int64_t)x1 * n2 + x2) * n3 + x3] : 0.f;
            Vs[MODE == LCPAO_BF16X3 ? tid : perm32(tid)] = v;
        }
        __syncthreads();
        if (has0) supertile<MODE>(Ahi, Alo, A32, Vs, si0, sj0, l16, q, acc0);
        if (has1) supertile<MODE>(Ahi, Alo, A32, Vs, si1, sj1, l16, q, acc1);
    }
    // lane holds D[4q + v][l16] of each 16x16 tile
#pragma unroll
    for (int h = 0; h < 2; h++) {
        if (h == 0 ? !has0 : !has1) continue;
        const int st = it.st0 + wave + h * HB_WAVES;
        float *out = part + (B.part_off + st) * (int64_t)(ST * ST);
        const f32x4 *acc = h == 0 ? acc0 : acc1;
#pragma unroll
        for (int t = 0; t < 16; t++) {
            const int rt = t >> 2, ct = t & 3;
#pragma unroll
            for (int v = 0; v < 4; v++) out[(rt * 16 + 4 * q + v) * ST + ct * 16 + l16] = acc[t][v];
        }
    }
}

// per computed pair: FP64 sum of its C_b sub-blocks over the blocks.  The
// pair's contribution records are staged in LDS so that every thread issues
// independent loads of the partial buffer.
constexpr int RED_BATCH = 256;
__global__ __launch_bounds__(256) void k_hblk_reduce(const int32_t *__restrict__ canon, int64_t ncanon,
                                                     const int32_t *__restrict__ coff,
                                                     const Contrib *__restrict__ contrib,
                                                     const BlockDev *__restrict__ blocks,
                                                     const PairDev *__restrict__ pairs, const float *__restrict__ part,
                                                     double gridvol, double *__restrict__ H) {
    __shared__ int64_t s_base[RED_BATCH];
    __shared__ int32_t s_r0[RED_BATCH], s_c0[RED_BATCH], s_tr[RED_BATCH], s_S[RED_BATCH];
    const int64_t ci = blockIdx.x;
    if (ci >= ncanon) return;
    const int p = canon[ci];
    const PairDev P = pairs[p];
    const int ne = P.noa * P.nob;
    const int k0 = coff[ci], k1 = coff[ci + 1];
    double acc[2] = {0.0, 0.0};   // elements threadIdx.x and threadIdx.x + 256 (ne <= 32*32 -> 4 per thread max)
    double acc2[2] = {0.0, 0.0};
    for (int kb = k0; kb < k1; kb += RED_BATCH) {
        const int nb = min(RED_BATCH, k1 - kb);
        __syncthreads();
        if ((int)threadIdx.x < nb) {
            const Contrib C = contrib[kb + threadIdx.x];
            const BlockDev B = blocks[C.block];
            s_base[threadIdx.x] = B.part_off * (int64_t)(ST * ST);
            s_r0[threadIdx.x] = C.r0;
            s_c0[threadIdx.x] = C.c0;
            s_tr[threadIdx.x] = C.trans;
            s_S[threadIdx.x] = B.rows / ST;
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 4; h++) {
            const int e = threadIdx.x + h * 256;
            if (e >= ne) break;
            const int i = e / P.nob, j = e - i * P.nob;
            double s = 0.0;
#pragma unroll 4
            for (int k = 0; k < nb; k++) {
                int r = s_tr[k] ? s_r0[k] + j : s_r0[k] + i;
                int cc = s_tr[k] ? s_c0[k] + i : s_c0[k] + j;
                if (r > cc) { const int t = r; r = cc; cc = t; }
                const int st = st_index(r / ST, cc / ST, s_S[k]);
                s += (double)part[s_base[k] + (int64_t)st * (ST * ST) + (r % ST) * ST + (cc % ST)];
            }
            if (h < 2) acc[h] += s; else acc2[h - 2] += s;
        }
    }
#pragma unroll
    for (int h = 0; h < 4; h++) {
        const int e = threadIdx.x + h * 256;
        if (e >= ne) break;
        const int i = e / P.nob, j = e - i * P.nob;
        const double s = (h < 2 ? acc[h] : acc2[h - 2]) * gridvol;
        H[P.hoff + e] = s;
        if (P.partner >= 0 && P.partner != p) H[pairs[P.partner].hoff + (int64_t)j * P.noa + i] = s;
    }
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

}  // namespace

void plan_free(BlockPlan *p) {
    if (!p) return;
    void *ptrs[] = {p->d_blocks, p->d_items, p->d_sph_block, p->d_sph_atom, p->d_sph_row0, p->d_sph_map,
                    p->d_contrib, p->d_contrib_off, p->d_hi, p->d_lo, p->d_t32, p->d_part, p->d_thi, p->d_tlo,
                    p->d_w32, p->d_contrib_pair};
    for (void *x : ptrs)
        if (x) (void)hipFree(x);
    delete p;
}

void invalidate_split_tables(lcpao_ctx *ctx) {
    if (ctx->plan) { ctx->plan->bf16_ready = false; ctx->plan->f32_ready = false; }
    if (ctx->d_tab32) { (void)hipFree(ctx->d_tab32); ctx->d_tab32 = nullptr; }
    for (int x = 0; x < 3; x++)
        if (ctx->d_dtab32[x]) { (void)hipFree(ctx->d_dtab32[x]); ctx->d_dtab32[x] = nullptr; }
}

int plan_build(lcpao_ctx *ctx) {
    if (ctx->plan) return LCPAO_OK;
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
    int64_t toff = 0, poff = 0;
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
        P->blocks[b] = {toff, poff, R, b1 * BLK, b2 * BLK, b3 * BLK};
        toff += (int64_t)R * BLK_PTS;
        poff += nst;
        for (int st0 = 0; st0 < nst; st0 += HB_WAVES * ST_PER_WAVE)
            P->items.push_back({(int32_t)b, st0, std::min(HB_WAVES * ST_PER_WAVE, nst - st0), 0});
        sph_first[b + 1] = sph_first[b] + (int64_t)sph[b].size();
    }
    P->tile_elems = toff;
    P->part_tiles = poff;
    P->nsph = sph_first[NB];
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
    struct CRec { int32_t cpos; Contrib c; };
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
                            int32_t pc = -1, trans = 0;
                            if (pxy >= 0 && canon_pos[pxy] >= 0) pc = pxy;
                            else if (pyx >= 0 && canon_pos[pyx] >= 0) { pc = pyx; trans = 1; }
                            if (pc < 0) continue;
                            recs[w].push_back({canon_pos[pc], {(int32_t)b, rows0[b][x], rows0[b][y], trans}});
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
    std::stable_sort(all.begin(), all.end(), [](const CRec &x, const CRec &y) {
        return x.cpos != y.cpos ? x.cpos < y.cpos : x.c.block < y.c.block;
    });
    std::vector<Contrib> contrib(all.size());
    std::vector<int32_t> cpair(all.size());
    std::vector<int32_t> coff(ctx->canon.size() + 1, 0);
    for (size_t i = 0; i < all.size(); i++) {
        contrib[i] = all[i].c;
        cpair[i] = ctx->canon[all[i].cpos];
        coff[all[i].cpos + 1]++;
    }
    P->ncontrib = (int64_t)all.size();
    for (size_t c = 0; c < ctx->canon.size(); c++) coff[c + 1] += coff[c];
    int rc;
    ctx->plan = P;
    if ((rc = up(ctx, &P->d_blocks, P->blocks)) || (rc = up(ctx, &P->d_items, P->items)) ||
        (rc = up(ctx, &P->d_sph_block, sb)) || (rc = up(ctx, &P->d_sph_atom, sa)) || (rc = up(ctx, &P->d_sph_row0, sr)) ||
        (rc = up(ctx, &P->d_sph_map, smap)) || (rc = up(ctx, &P->d_contrib, contrib)) || (rc = up(ctx, &P->d_contrib_off, coff)) ||
        (rc = up(ctx, &P->d_contrib_pair, cpair))) {
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
    if (!P->d_part) SPLIT_TRY(ctx, hipMalloc(&P->d_part, std::max<int64_t>(1, P->part_tiles) * ST * ST * sizeof(float)));
    SPLIT_TRY(ctx, hipStreamSynchronize(ctx->stream));
    (bf ? P->bf16_ready : P->f32_ready) = true;
    return LCPAO_OK;
}

int split_hamiltonian(lcpao_ctx *ctx, int nch, const double *const *d_V, double *const *d_H, int precision) {
    int rc = split_tables(ctx, precision);
    if (rc) return rc;
    BlockPlan *P = ctx->plan;
    const bool bf = precision == LCPAO_BF16X3;
    const size_t lds = bf ? (size_t)P->max_rows * LROW * 2 * sizeof(uint16_t) + CHUNK * sizeof(float)
                          : (size_t)P->max_rows * LROW32 * sizeof(float) + CHUNK * sizeof(float);
    const unsigned nitems = (unsigned)P->items.size();
    for (int c = 0; c < nch; c++) {
        if (!d_V[c] || !d_H[c]) { ctx->err = "null channel pointer"; return LCPAO_EINVAL; }
        if (nitems) {
            KTimer kt(ctx, KID_HAM);
            if (bf)
                hipLaunchKernelGGL(k_hblk<LCPAO_BF16X3>, dim3(nitems), dim3(64 * HB_WAVES), lds, ctx->stream, P->d_items,
                                   P->d_blocks, P->d_hi, P->d_lo, P->d_t32, d_V[c], ctx->ngrid[0], ctx->ngrid[1],
                                   ctx->ngrid[2], P->d_part);
            else
                hipLaunchKernelGGL(k_hblk<LCPAO_FP32>, dim3(nitems), dim3(64 * HB_WAVES), lds, ctx->stream, P->d_items,
                                   P->d_blocks, P->d_hi, P->d_lo, P->d_t32, d_V[c], ctx->ngrid[0], ctx->ngrid[1],
                                   ctx->ngrid[2], P->d_part);
        }
        if (ctx->ncanon > 0) {
            KTimer kt(ctx, KID_HRED);
            hipLaunchKernelGGL(k_hblk_reduce, dim3((unsigned)ctx->ncanon), dim3(256), 0, ctx->stream, ctx->d_hcanon,
                               ctx->ncanon, P->d_contrib_off, P->d_contrib, P->d_blocks, ctx->d_pairs, P->d_part,
                               ctx->gridvol, d_H[c]);
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
//   assembled per call from the folded pair blocks.
// ----------------------------------------------------------------------------
// One workgroup per computed pair: its folded density block is read once and
// written into every block super-tile where the pair's spheres overlap.
__global__ __launch_bounds__(256) void k_wasm(const int32_t *__restrict__ canon, int64_t ncanon,
                                              const int32_t *__restrict__ coff, const Contrib *__restrict__ contrib,
                                              const BlockDev *__restrict__ blocks, const PairDev *__restrict__ pairs,
                                              const double *__restrict__ Deff, float *__restrict__ w32) {
    const int64_t ci = blockIdx.x;
    if (ci >= ncanon) return;
    const PairDev P = pairs[canon[ci]];
    const int ne = P.noa * P.nob;
    const double *D = Deff + P.hoff;
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
        const int i = e / P.nob, j = e - i * P.nob;
        const double d = D[e];
        const double dsym = (P.noa == P.nob) ? d + D[j * P.nob + i] : d;  // used only for on-site blocks
        for (int k = coff[ci]; k < coff[ci + 1]; k++) {
            const Contrib C = contrib[k];
            const BlockDev B = blocks[C.block];
            const int S = B.rows / ST;
            const bool same = C.r0 == C.c0;
            const double val = same ? dsym : d;
            const int r = C.trans ? C.r0 + j : C.r0 + i;
            const int c = C.trans ? C.c0 + i : C.c0 + j;
            const int sr = r / ST, sc = c / ST;
            const int64_t base = B.part_off * (int64_t)(ST * ST);
            if (sr == sc) {
                const int64_t st = base + (int64_t)st_index(sr, sr, S) * (ST * ST);
                w32[st + (r % ST) * ST + (c % ST)] = (float)(0.5 * val);
                if (!same) w32[st + (c % ST) * ST + (r % ST)] = (float)(0.5 * val);
            } else if (r < c) {
                w32[base + (int64_t)st_index(sr, sc, S) * (ST * ST) + (r % ST) * ST + (c % ST)] = (float)val;
            }
        }
    }
}

constexpr int TLROW_PAD = 8;  // LDS row padding of the transposed chunk (bf16 elements)

__global__ __launch_bounds__(64 * HB_WAVES) void k_rblk(const HItem *__restrict__ items, const BlockDev *__restrict__ blocks,
                                                       const uint16_t *__restrict__ thi, const uint16_t *__restrict__ tlo,
                                                       const float *__restrict__ w32, int n1, int n2, int n3,
                                                       double *__restrict__ rho) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const HItem it = items[blockIdx.x];
    const BlockDev B = blocks[it.block];
    const int R = B.rows, S = R / ST, LR = R + TLROW_PAD;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, l16 = lane & 15, q = lane >> 4;
    uint16_t *Th = reinterpret_cast<uint16_t *>(lds_raw);  // [point][row]
    uint16_t *Tl = Th + CHUNK * LR;
    float *rc = reinterpret_cast<float *>(Tl + CHUNK * LR);  // 32 point sums
    // W fragments of this wave's super-tiles stay in registers for the whole block
    bf16x8 wh[2][4][2], wl[2][4][2];
    int si[2] = {0, 0}, sj[2] = {0, 0};
    bool has[2];
    for (int h = 0; h < 2; h++) {
        has[h] = wave + h * HB_WAVES < it.nst;
        if (!has[h]) continue;
        const int st = it.st0 + wave + h * HB_WAVES;
        st_decode(st, S, si[h], sj[h]);
        const int64_t base = (B.part_off + st) * (int64_t)(ST * ST);
#pragma unroll
        for (int rt = 0; rt < 4; rt++)
#pragma unroll
            for (int ks = 0; ks < 2; ks++) {
                const int64_t idx = base + (rt * 16 + l16) * ST + ks * 32 + q * 8;
                const f32x8 w = *reinterpret_cast<const f32x8 *>(&w32[idx]);
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const __bf16 hb = (__bf16)w[e];
                    wh[h][rt][ks][e] = hb;
                    wl[h][rt][ks][e] = (__bf16)(w[e] - (float)hb);
                }
            }
    }
    for (int c = 0; c < NCHUNK; c++) {
        __syncthreads();
        const uint16_t *sh = thi + B.tile_off + (int64_t)c * CHUNK * R;
        const uint16_t *sl = tlo + B.tile_off + (int64_t)c * CHUNK * R;
        for (int e = tid * 8; e < CHUNK * R; e += 64 * HB_WAVES * 8) {
            const int pt = e / R, row = e % R;
            *reinterpret_cast<u16x8 *>(&Th[pt * LR + row]) = *reinterpret_cast<const u16x8 *>(&sh[e]);
            *reinterpret_cast<u16x8 *>(&Tl[pt * LR + row]) = *reinterpret_cast<const u16x8 *>(&sl[e]);
        }
        if (tid < CHUNK) rc[tid] = 0.f;
        __syncthreads();
        for (int h = 0; h < 2; h++) {
            if (!has[h]) continue;
            float part[2] = {0.f, 0.f};
#pragma unroll
            for (int pt = 0; pt < 2; pt++) {
                const int point = pt * 16 + l16;
                bf16x8 bh[2], bl[2];
#pragma unroll
                for (int ks = 0; ks < 2; ks++) {
                    const int row = sj[h] * ST + ks * 32 + q * 8;
                    bh[ks] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(&Th[point * LR + row]));
                    bl[ks] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(&Tl[point * LR + row]));
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
                    // acc[v] = Q[row si*64 + rt*16 + 4q + v][point]
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int row = si[h] * ST + rt * 16 + 4 * q + v;
                        const float ph = (float)__builtin_bit_cast(__bf16, Th[point * LR + row]);
                        const float pl = (float)__builtin_bit_cast(__bf16, Tl[point * LR + row]);
                        part[pt] = fmaf(ph + pl, acc[v], part[pt]);
                    }
                }
            }
#pragma unroll
            for (int pt = 0; pt < 2; pt++) {
                float x = part[pt];
                x += __shfl_xor(x, 16, 64);
                x += __shfl_xor(x, 32, 64);
                if (q == 0) atomicAdd(&rc[pt * 16 + l16], x);
            }
        }
        __syncthreads();
        if (tid < CHUNK) {
            const int k = c * CHUNK + tid;
            const int x1 = B.g1 + k / (BLK * BLK), x2 = B.g2 + (k / BLK) % BLK, x3 = B.g3 + k % BLK;
            if (x1 < n1 && x2 < n2 && x3 < n3) atomicAdd(&rho[((int64_t)x1 * n2 + x2) * n3 + x3], (double)rc[tid]);
        }
    }
}

int split_density(lcpao_ctx *ctx, int nspin, const double *const *d_DM, int64_t dm_offset, double *const *d_rho) {
    int rc = split_tables(ctx, LCPAO_BF16X3);
    if (rc) return rc;
    BlockPlan *P = ctx->plan;
    if (!P->d_w32) {
        const size_t bytes = std::max<int64_t>(1, P->part_tiles) * ST * ST * sizeof(float);
        SPLIT_TRY(ctx, hipMalloc(&P->d_w32, bytes));
        SPLIT_TRY(ctx, hipMemsetAsync(P->d_w32, 0, bytes, ctx->stream));
    }
    const size_t lds = (size_t)CHUNK * (P->max_rows + TLROW_PAD) * 2 * sizeof(uint16_t) + CHUNK * sizeof(float);
    const unsigned nitems = (unsigned)P->items.size();
    for (int s = 0; s < nspin; s++) {
        if (!d_DM[s] || !d_rho[s]) { ctx->err = "null spin pointer"; return LCPAO_EINVAL; }
        const double *D = d_DM[s] + dm_offset;
        if (ctx->ncanon > 0) {
            KTimer kt(ctx, KID_DEFF);
            launch_deff_f64(ctx, D);
            hipLaunchKernelGGL(k_wasm, dim3((unsigned)ctx->ncanon), dim3(256), 0, ctx->stream, ctx->d_hcanon, ctx->ncanon,
                               P->d_contrib_off, P->d_contrib, P->d_blocks, ctx->d_pairs, ctx->d_deff, P->d_w32);
        }
        SPLIT_TRY(ctx, hipMemsetAsync(d_rho[s], 0, (size_t)ctx->nn * sizeof(double), ctx->stream));
        if (nitems) {
            KTimer kt(ctx, KID_RHO);
            hipLaunchKernelGGL(k_rblk, dim3(nitems), dim3(64 * HB_WAVES), lds, ctx->stream, P->d_items, P->d_blocks,
                               P->d_thi, P->d_tlo, P->d_w32, ctx->ngrid[0], ctx->ngrid[1], ctx->ngrid[2], d_rho[s]);
        }
        SPLIT_TRY(ctx, hipGetLastError());
    }
    SPLIT_TRY(ctx, hipStreamSynchronize(ctx->stream));
    collect_timings(ctx);
    return LCPAO_OK;
}

}  // namespace lcpao
