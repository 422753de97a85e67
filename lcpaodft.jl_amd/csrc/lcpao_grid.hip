// lcpao_grid.hip -- MI355X (gfx950) kernels and C ABI for the LCPAO
// real-space grid contractions: grid Hamiltonian, density and Force3.
//
// Work decomposition of the pair-list kernels (on the reference's atom-pair
// overlap lists; the block forms are in split.hip):
//  * H  : pair-major.  Each "canonical" pair P (the one of P=(A,B,Rn) and its
//         transpose P'=(B,A,-Rn) with the lower index, when both are local)
//         is cut into units of <= 512 overlap points.  One wave per unit
//         gathers phi_A and V*phi_B straight from the tables as the operands
//         of FP64 MFMAs (4x4 tiles of the 4x4x4 form up to 20 orbitals per
//         atom; 2x2 tiles of 16x16, one 32x32 chunk per unit, up to 32;
//         32x32 chunks in a loop beyond); units write partial blocks that a
//         second kernel sums per pair in a fixed order (deterministic),
//         scales by GridVol and writes H[P] and H[P']^T.
//  * rho: atom-owned.  Work unit = (home atom A, chunk of grid-list entries
//         of A).  Each wave walks the (pair, sub-range) segments of A's
//         canonical pairs inside the chunk; a lane owns one overlap point,
//         the pair's folded density block D_P + D_P'^T is wave-uniform and
//         read through the scalar path.  Per-wave LDS accumulators are
//         summed in a fixed order; a gather kernel then sums the entries of
//         every grid point (CSR by grid index) -- no atomics anywhere.
//  * Force3: pair-major units of <= 256 overlap points over all local
//         pairs, lane = overlap point, per-unit partial sums reduced per
//         atom in a fixed order.
// Units are launched in order of the grid index of their first point, so
// the units in flight touch a few grid planes and the tables they gather
// stay in the Infinity Cache, dealt to the XCDs in contiguous ranges so
// that each XCD's L2 serves its own band of planes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "split.hpp"

namespace {
std::string g_create_err;
}  // namespace

// Scratch of the allocation-free FP64 Force3 entry point
// (lcpao_grid_force3_dev_noalloc): the per-unit partial sums (3 doubles per
// Force3 unit) followed by the per-atom sums (3 doubles per atom).  Every GPU
// that loads the library holds one copy of this array until the process
// ends; calls on one GPU share it.
#define NOALLOC_SCRATCH_DOUBLES (16 * 1024 * 1024)  // 128 MiB
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {
// Device address of `scratch` on each device, looked up once per device by
// lcpao_grid_create with that device current.  The host-side name `scratch`
// is not a device address and is never passed to a kernel or copy.
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
hipError_t g_scratch_rc[NOALLOC_MAX_DEVICES] = {};
bool g_scratch_done[NOALLOC_MAX_DEVICES] = {};

void scratch_setup(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_done[device]) return;
    void *p = nullptr;
    g_scratch_rc[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_ptr[device] = g_scratch_rc[device] == hipSuccess ? static_cast<double *>(p) : nullptr;
    g_scratch_done[device] = true;
}
}  // namespace

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

typedef double d4_t __attribute__((ext_vector_type(4)));

// Hamiltonian unit kernel on the FP64 matrix cores (v_mfma_f64_16x16x4f64).
// One wave per unit.  The NO_A x NO_B block is computed in 32 x 32 chunks,
// each held as 2x2 MFMA tiles: one chunk for up to 32 orbitals per atom;
// larger bases loop over the chunks and read the unit's points again for
// each.  Lane l supplies A[l%16][l/16] = phi_A row of point l/16 and
// B[l/16][l%16] = V*phi_B column, gathered straight from the tables; the 4
// points of a step are consecutive overlap points.
__device__ __forceinline__ void ham_chunk64(const double *__restrict__ tabA, const double *__restrict__ tabB,
                                            const int32_t *__restrict__ L1, const int32_t *__restrict__ L2,
                                            const int32_t *__restrict__ glA, const double *__restrict__ V, int n,
                                            int noa, int nob, int r0, int c0, int lane, double *__restrict__ out) {
    const int r16 = lane & 15, kq = lane >> 4;
    d4_t c00 = {0, 0, 0, 0}, c01 = c00, c10 = c00, c11 = c00;
    const bool ra0 = r0 + r16 < noa, ra1 = r0 + 16 + r16 < noa, rb0 = c0 + r16 < nob, rb1 = c0 + 16 + r16 < nob;
    for (int g0 = 0; g0 < n; g0 += 64) {
        // indices and potential of 64 points, one per lane
        // (unconditional loads from clamped indices, zeroed after, so that
        // the gathers issue together instead of one branch each)
        const int gi = g0 + lane;
        const bool gok = gi < n;
        const int gc = gok ? gi : 0;
        const int t1g = L1[gc];
        const int t2g = L2[gc];
        const double vv = V[glA[t1g]];
        const double vg = gok ? vv : 0.0;
        const int nstep = min(16, (n - g0 + 3) >> 2);
        const int ia0 = ra0 ? r16 : 0, ia1 = ra1 ? 16 + r16 : 0, ib0 = rb0 ? r16 : 0, ib1 = rb1 ? 16 + r16 : 0;
#pragma unroll 4
        for (int st = 0; st < nstep; st++) {
            const int src = 4 * st + kq;
            const int t1 = __shfl(t1g, src, 64);
            const int t2 = __shfl(t2g, src, 64);
            const double v = __shfl(vg, src, 64);  // 0 for points past the end
            const double *ca = tabA + (int64_t)t1 * noa + r0, *cb = tabB + (int64_t)t2 * nob + c0;
            const double xa0 = ca[ia0], xa1 = ca[ia1], xb0 = cb[ib0], xb1 = cb[ib1];
            const double a0 = ra0 ? xa0 : 0.0, a1 = ra1 ? xa1 : 0.0;
            const double b0 = rb0 ? xb0 * v : 0.0, b1 = rb1 ? xb1 * v : 0.0;
            c00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, c00, 0, 0, 0);
            c01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, c01, 0, 0, 0);
            c10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, c10, 0, 0, 0);
            c11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, c11, 0, 0, 0);
        }
    }
    // v_mfma_f64_16x16x4f64 keeps D[kq + 4v][r16] in lane (16 kq + r16), register v
    const d4_t *cs[4] = {&c00, &c01, &c10, &c11};
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int rt = t >> 1, ct = t & 1;
        const int col = c0 + ct * 16 + r16;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = r0 + rt * 16 + kq + 4 * v;
            if (row < noa && col < nob) out[row * nob + col] = (*cs[t])[v];
        }
    }
}

// Hamiltonian unit kernel on the 4x4x4 FP64 matrix-core form
// (v_mfma_f64_4x4x4f64: four independent 4x4x4 products per instruction;
// on gfx950 it issues 512 FLOP per ~18 cycles against 2048 per ~105 for
// the 16x16x4 form, 1.45x the rate).  One wave per unit; the NO_A x NO_B
// block as NTA x NTB tiles of 4 x 4 (19 orbitals: 5 x 5 tiles, 90 % useful
// against 35 % for 32 x 32), each tile accumulated in one register per lane.
// Block b of an instruction takes 4 of every 16 overlap points: lane
// l = 16 k + 4 b + r supplies A[r][k] = phi_A[4 ti + r](point 4 b + k) and
// B[k][r] = V phi_B[4 tj + r](point 4 b + k), and D[m][n] of block b lands
// in lane 16 m + 4 b + n; the four blocks' partial tiles are added with
// lane swaps at the end of the unit (fixed order).  Splitting a unit's tile
// rows over two waves (eight waves per SIMD instead of four, each loading
// the V phi_B operands again) ran at 10.5 ms against 7.6 ms on GaAs216.
template <int NTA, int NTB>
__device__ __forceinline__ void ham_unit_m4(const double *__restrict__ tabA, const double *__restrict__ tabB,
                                            const int32_t *__restrict__ L1, const int32_t *__restrict__ L2,
                                            const int32_t *__restrict__ glA, const double *__restrict__ V, int n,
                                            int noa, int nob, int lane, double *__restrict__ out) {
    const int r = lane & 3, pb = ((lane >> 2) & 3) * 4 + (lane >> 4);  // row/column in a tile, point in a group of 16
    double acc[NTA][NTB];
#pragma unroll
    for (int ti = 0; ti < NTA; ti++)
#pragma unroll
        for (int tj = 0; tj < NTB; tj++) acc[ti][tj] = 0.0;
    for (int g0 = 0; g0 < n; g0 += 64) {
        // indices and potential of 64 points, one per lane; the loads are
        // unconditional (a lane past the end reads point 0 and is zeroed
        // after), so that they issue together instead of one branch each
        const int gi = g0 + lane;
        const bool gok = gi < n;
        const int gc = gok ? gi : 0;
        const int t1g = L1[gc];
        const int t2g = L2[gc];
        const double vv = V[glA[t1g]];
        const double vg = gok ? vv : 0.0;
        const int ngrp = min(4, (n - g0 + 15) >> 4);
        for (int st = 0; st < ngrp; st++) {
            const int src = 16 * st + pb;
            const int t1 = __shfl(t1g, src, 64);
            const int t2 = __shfl(t2g, src, 64);
            const double v = __shfl(vg, src, 64);  // 0 past the end
            const double *ca = tabA + (int64_t)t1 * noa, *cb = tabB + (int64_t)t2 * nob;
            double a[NTA], b[NTB];
#pragma unroll
            for (int ti = 0; ti < NTA; ti++) {
                const int i = 4 * ti + r;
                const double x = ca[i < noa ? i : noa - 1];
                a[ti] = i < noa ? x : 0.0;
            }
#pragma unroll
            for (int tj = 0; tj < NTB; tj++) {
                const int j = 4 * tj + r;
                const double x = cb[j < nob ? j : nob - 1];
                b[tj] = j < nob ? x * v : 0.0;
            }
#pragma unroll
            for (int ti = 0; ti < NTA; ti++)
#pragma unroll
                for (int tj = 0; tj < NTB; tj++)
                    acc[ti][tj] = __builtin_amdgcn_mfma_f64_4x4x4f64(a[ti], b[tj], acc[ti][tj], 0, 0, 0);
        }
    }
    // lane 16 m + 4 b + n holds block b's D[m][n]: add the blocks (b = lane bits 2-3)
    const int m = lane >> 4, nn = lane & 3;
#pragma unroll
    for (int ti = 0; ti < NTA; ti++)
#pragma unroll
        for (int tj = 0; tj < NTB; tj++) {
            double x = acc[ti][tj];
            x += __shfl_xor(x, 4, 64);
            x += __shfl_xor(x, 8, 64);
            const int row = 4 * ti + m, col = 4 * tj + nn;
            if (((lane >> 2) & 3) == 0 && row < noa && col < nob) out[row * nob + col] = x;
        }
}

// the 4x4x4 form for pairs of up to 20 orbitals per atom (4 or 5 tiles a side)
__global__ __launch_bounds__(64) void k_ham_m4(const HUnit *__restrict__ units, const int32_t *__restrict__ order,
                                               int32_t nunits, const PairDev *__restrict__ pairs,
                                               const double *__restrict__ tab, const int32_t *__restrict__ olg1,
                                               const int32_t *__restrict__ olg2, const int32_t *__restrict__ glist,
                                               const double *__restrict__ V, double *__restrict__ part) {
    const int b = blockIdx.x;
    if (b >= nunits || order[b] < 0) return;
    const HUnit u = units[order[b]];
    const PairDev P = pairs[__builtin_amdgcn_readfirstlane(u.pair)];
    const int noa = __builtin_amdgcn_readfirstlane(P.noa), nob = __builtin_amdgcn_readfirstlane(P.nob);
    const int lane = threadIdx.x;
    const double *tabA = tab + P.tab_a, *tabB = tab + P.tab_b;
    const int32_t *L1 = olg1 + P.olg + u.lo, *L2 = olg2 + P.olg + u.lo;
    const int32_t *glA = glist + P.gl_a;
    double *out = part + u.part;
    const int ta = (noa + 3) >> 2, tb = (nob + 3) >> 2;
    if (ta <= 4 && tb <= 4) ham_unit_m4<4, 4>(tabA, tabB, L1, L2, glA, V, u.n, noa, nob, lane, out);
    else if (ta <= 4) ham_unit_m4<4, 5>(tabA, tabB, L1, L2, glA, V, u.n, noa, nob, lane, out);
    else if (tb <= 4) ham_unit_m4<5, 4>(tabA, tabB, L1, L2, glA, V, u.n, noa, nob, lane, out);
    else ham_unit_m4<5, 5>(tabA, tabB, L1, L2, glA, V, u.n, noa, nob, lane, out);
}
constexpr int M4_NO = 20;  // orbitals per atom the 4x4x4 form holds

// CHUNKED: any orbital count (a loop over the 32 x 32 chunks); else every
// atom holds at most 32 orbitals and the block is one chunk -- a kernel of
// its own, since the loop's registers (42 VGPRs against 24) cost it two
// waves per SIMD of the occupancy that hides its gathers (GaAs216: 10.3 ms
// against 9.3 ms).
template <bool CHUNKED>
__global__ __launch_bounds__(64) void k_ham_mfma64(const HUnit *__restrict__ units, const int32_t *__restrict__ order,
                                                   int32_t nunits, const PairDev *__restrict__ pairs,
                                                   const double *__restrict__ tab, const int32_t *__restrict__ olg1,
                                                   const int32_t *__restrict__ olg2, const int32_t *__restrict__ glist,
                                                   const double *__restrict__ V, double *__restrict__ part) {
    const int b = blockIdx.x;
    if (b >= nunits || order[b] < 0) return;
    const HUnit u = units[order[b]];
    const PairDev P = pairs[__builtin_amdgcn_readfirstlane(u.pair)];
    const int noa = __builtin_amdgcn_readfirstlane(P.noa), nob = __builtin_amdgcn_readfirstlane(P.nob);
    const int lane = threadIdx.x;
    const double *tabA = tab + P.tab_a, *tabB = tab + P.tab_b;
    const int32_t *L1 = olg1 + P.olg + u.lo, *L2 = olg2 + P.olg + u.lo;
    const int32_t *glA = glist + P.gl_a;
    double *out = part + u.part;
    if constexpr (!CHUNKED) {
        ham_chunk64(tabA, tabB, L1, L2, glA, V, u.n, noa, nob, 0, 0, lane, out);
    } else {
        for (int r0 = 0; r0 < noa; r0 += 32)
            for (int c0 = 0; c0 < nob; c0 += 32) ham_chunk64(tabA, tabB, L1, L2, glA, V, u.n, noa, nob, r0, c0, lane, out);
    }
}

__global__ void k_fill_zero(int64_t n, double *__restrict__ x) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

// Sums the unit partials of every computed pair in unit order, scales by
// GridVol and writes H[P] (and H[P']^T for a folded transpose partner).
__global__ void k_ham_reduce(const int32_t *__restrict__ canon, const int32_t *__restrict__ ub, int64_t ncanon,
                             const HUnit *__restrict__ units, const PairDev *__restrict__ pairs,
                             const double *__restrict__ part, double gridvol, double *__restrict__ H) {
    const int64_t c = blockIdx.x;
    if (c >= ncanon) return;
    const int p = canon[c];
    const PairDev P = pairs[p];
    const int ne = P.noa * P.nob;
    const int u0 = ub[c], u1 = ub[c + 1];
    // a pair's units are consecutive and so are their partial blocks (a pair
    // without overlap points has none: its H block is zero)
    const int nu = u1 - u0;
    const double *pp = part + (nu > 0 ? units[u0].part : 0);
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
        double s = 0.0;
        int u = 0;
        for (; u + 4 <= nu; u += 4) {  // four loads in flight, added in unit order
            const double v0 = pp[(int64_t)u * ne + e], v1 = pp[(int64_t)(u + 1) * ne + e];
            const double v2 = pp[(int64_t)(u + 2) * ne + e], v3 = pp[(int64_t)(u + 3) * ne + e];
            s += v0;
            s += v1;
            s += v2;
            s += v3;
        }
        for (; u < nu; u++) s += pp[(int64_t)u * ne + e];
        s *= gridvol;
        H[P.hoff + e] = s;
        if (P.partner >= 0 && P.partner != p) {
            const PairDev Q = pairs[P.partner];
            const int i = e / P.nob, j = e - i * P.nob;
            H[Q.hoff + (int64_t)j * P.noa + i] = s;
        }
    }
}

// Folded density block of each computed pair: D_P + D_P'^T (or D_P alone).
// sym_self != 0: an on-site pair (its own transpose) gets D_P + D_P^T, the
// form the block-dense density reads (one load per W element).
template <typename T>
__device__ __forceinline__ void deff_pair(int64_t c, const int32_t *__restrict__ canon,
                                          const PairDev *__restrict__ pairs, const double *__restrict__ D,
                                          T *__restrict__ Deff, int sym_self) {
    const int p = canon[c];
    const PairDev P = pairs[p];
    const int ne = P.noa * P.nob;
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
        double v = D[P.hoff + e];
        if (P.partner >= 0 && P.partner != p) {
            const PairDev Q = pairs[P.partner];
            const int i = e / P.nob, j = e - i * P.nob;
            v += D[Q.hoff + (int64_t)j * P.noa + i];
        } else if (sym_self && P.partner == p) {
            const int i = e / P.nob, j = e - i * P.nob;
            v += D[P.hoff + (int64_t)j * P.nob + i];
        }
        Deff[P.hoff + e] = (T)v;
    }
}

template <typename T>
__global__ void k_deff(const int32_t *__restrict__ canon, int64_t ncanon, const PairDev *__restrict__ pairs,
                       const double *__restrict__ D, T *__restrict__ Deff, int sym_self = 0) {
    const int64_t c = blockIdx.x;
    if (c >= ncanon) return;
    deff_pair<T>(c, canon, pairs, D, Deff, sym_self);
}

// the FP32 folds (on-site blocks symmetrised) of every structure of a batch
// in one launch over the computed pairs of all structures (a flat grid)
__global__ void k_deff_batch(const lcpao::BatchFold *__restrict__ fold, const int32_t *__restrict__ off, int nstruct) {
    const int w = blockIdx.x;
    const int st = lcpao::batch_struct_of(off, nstruct, w);
    const lcpao::BatchFold F = fold[st];
    const int64_t c = w - off[st];
    if (c >= F.ncanon) return;
    deff_pair<float>(c, F.canon, F.pairs, F.D, F.Deff, 1);
}

// ---- density: point work of one segment -------------------------------------
// A lane owns one overlap point (two per iteration).  D (the folded pair
// block) is wave-uniform: with NA, NB known at compile time its loads are
// batched scalar loads whose values feed v_fma_f64 as SGPR operands.
template <typename T, int NA, int NB>
__device__ __forceinline__ void rho_points(const T *__restrict__ D, const T *__restrict__ tabA,
                                           const T *__restrict__ tabB, const int32_t *__restrict__ L1,
                                           const int32_t *__restrict__ L2, int lo, int hi, int t0, int lane,
                                           double *__restrict__ racc) {
    for (int base = lo; base < hi; base += 128) {
        const int i0 = base + lane, i1 = base + 64 + lane;
        const bool ok0 = i0 < hi, ok1 = i1 < hi;
        const int t1a = L1[ok0 ? i0 : lo], t2a = L2[ok0 ? i0 : lo];
        const int t1b = L1[ok1 ? i1 : lo], t2b = L2[ok1 ? i1 : lo];
        const T *cb0 = tabB + (int64_t)t2a * NB, *cb1 = tabB + (int64_t)t2b * NB;
        const T *ca0 = tabA + (int64_t)t1a * NA, *ca1 = tabA + (int64_t)t1b * NA;
        T b0[NB], b1[NB];
#pragma unroll
        for (int j = 0; j < NB; j++) { b0[j] = cb0[j]; b1[j] = cb1[j]; }
        T z0 = 0, z1 = 0;
#pragma unroll
        for (int i = 0; i < NA; i++) {
            T y0 = 0, y1 = 0;
#pragma unroll
            for (int j = 0; j < NB; j++) {
                const T d = D[i * NB + j];
                y0 = fma(d, b0[j], y0);
                y1 = fma(d, b1[j], y1);
            }
            z0 = fma(ca0[i], y0, z0);
            z1 = fma(ca1[i], y1, z1);
        }
        if (ok0) racc[t1a - t0] += (double)z0;
        if (ok1) racc[t1b - t0] += (double)z1;
    }
}

// Runtime-size fallback (any NO <= MAXNO).
template <typename T>
__device__ void rho_points_any(const T *__restrict__ D, const T *__restrict__ tabA, const T *__restrict__ tabB,
                               const int32_t *__restrict__ L1, const int32_t *__restrict__ L2, int lo, int hi,
                               int t0, int lane, double *__restrict__ racc, int noa, int nob) {
    for (int base = lo; base < hi; base += 64) {
        const int idx = base + lane;
        const bool ok = idx < hi;
        const int t1 = L1[ok ? idx : lo], t2 = L2[ok ? idx : lo];
        const T *cb = tabB + (int64_t)t2 * nob, *ca = tabA + (int64_t)t1 * noa;
        T z = 0;
        for (int i = 0; i < noa; i++) {
            T y = 0;
            for (int j = 0; j < nob; j++) y = fma(D[i * nob + j], cb[j], y);
            z = fma(ca[i], y, z);
        }
        if (ok) racc[t1 - t0] += (double)z;
    }
}


// Density per (atom, chunk) unit; T = table/D element type (double or float),
// W waves per workgroup, CH grid-list entries per unit (the unit shape: W
// per-wave accumulators of CH doubles in LDS).
template <typename T, int W, int CH>
__global__ __launch_bounds__(64 * W) void k_rho(
    const RhoUnit *__restrict__ units, const int32_t *__restrict__ order, int32_t nunits,
    const Seg *__restrict__ segs, const PairDev *__restrict__ pairs, const T *__restrict__ tab,
    const int32_t *__restrict__ olg1, const int32_t *__restrict__ olg2, const T *__restrict__ Deff,
    const int64_t *__restrict__ gl_off, double *__restrict__ rhoA) {
    __shared__ double racc[W][CH];
    const int b = blockIdx.x;
    if (b >= nunits) return;
    const RhoUnit u = units[order[b]];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int t = threadIdx.x; t < W * CH; t += blockDim.x) (&racc[0][0])[t] = 0.0;
    __syncthreads();
    for (int s = u.seg_beg + wave; s < u.seg_end; s += W) {
        const Seg sg = segs[s];
        const PairDev P = pairs[__builtin_amdgcn_readfirstlane(sg.pair)];
        const int noa = __builtin_amdgcn_readfirstlane(P.noa), nob = __builtin_amdgcn_readfirstlane(P.nob);
        const T *D = Deff + P.hoff;
        const T *tabA = tab + P.tab_a, *tabB = tab + P.tab_b;
        const int32_t *L1 = olg1 + P.olg, *L2 = olg2 + P.olg;
#define RHO_CASE(A, B)                                                                                 \
    if (noa == A && nob == B) {                                                                        \
        rho_points<T, A, B>(D, tabA, tabB, L1, L2, sg.lo, sg.hi, u.t0, lane, racc[wave]);              \
        continue;                                                                                      \
    }
        NO_CASES(RHO_CASE)
#undef RHO_CASE
        rho_points_any<T>(D, tabA, tabB, L1, L2, sg.lo, sg.hi, u.t0, lane, racc[wave], noa, nob);
    }
    __syncthreads();
    double *out = rhoA + gl_off[u.atom] + u.t0;
    for (int t = threadIdx.x; t < u.nt; t += blockDim.x) {
        double acc = racc[0][t];
#pragma unroll
        for (int w = 1; w < W; w++) acc += racc[w][t];
        out[t] = acc;
    }
}

// ---- FP64 density on the matrix cores ---------------------------------------
// Per segment (pair P, overlap points of one chunk of A's list): for 16 points
// at a time, Y = D_P (NO_A x NO_B) x Phi_B (NO_B x 16) with v_mfma_f64_16x16x4f64
// (rows padded to 32, k to a multiple of 4): D is the A operand, held in
// VGPRs for the whole segment (lane l: D[rt*16 + l%16][4s + l/16]); the B
// operand is gathered from the tables (lane l: phi_B[point l%16][4s + l/16]).
// Lane l receives Y[rt*16 + l/16 + 4v][point l%16]; z = phi_A . Y is summed
// over v and over the four lane groups.
typedef double d4_m __attribute__((ext_vector_type(4)));

template <int KS>
__device__ __forceinline__ void rho_seg_m64(const double *__restrict__ D, const double *__restrict__ tabA,
                                            const double *__restrict__ tabB, const int32_t *__restrict__ L1,
                                            const int32_t *__restrict__ L2, int lo, int hi, int t0, int noa, int nob,
                                            int lane, double *__restrict__ racc) {
    const int r16 = lane & 15, kq = lane >> 4;
    double a[2][KS];
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
        for (int st = 0; st < KS; st++) {
            const int i = rt * 16 + r16, j = 4 * st + kq;
            a[rt][st] = (i < noa && j < nob) ? D[i * nob + j] : 0.0;
        }
    for (int base = lo; base < hi; base += 32) {
        int t1[2], t2[2];
        bool ok[2];
        double b[2][KS];
#pragma unroll
        for (int g = 0; g < 2; g++) {
            const int idx = base + g * 16 + r16;
            ok[g] = idx < hi;
            t1[g] = L1[ok[g] ? idx : lo];
            t2[g] = L2[ok[g] ? idx : lo];
            const double *cb = tabB + (int64_t)t2[g] * nob;
#pragma unroll
            for (int st = 0; st < KS; st++) {
                const int j = 4 * st + kq;
                b[g][st] = (ok[g] && j < nob) ? cb[j] : 0.0;
            }
        }
        d4_m y[2][2];
#pragma unroll
        for (int g = 0; g < 2; g++)
#pragma unroll
            for (int rt = 0; rt < 2; rt++) {
                y[g][rt] = d4_m{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int st = 0; st < KS; st++)
                    y[g][rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt][st], b[g][st], y[g][rt], 0, 0, 0);
            }
#pragma unroll
        for (int g = 0; g < 2; g++) {
            const double *ca = tabA + (int64_t)t1[g] * noa;
            double z = 0.0;
#pragma unroll
            for (int rt = 0; rt < 2; rt++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int i = rt * 16 + kq + 4 * v;
                    if (i < noa) z = fma(ca[i], y[g][rt][v], z);
                }
            z += __shfl_xor(z, 16, 64);
            z += __shfl_xor(z, 32, 64);
            if (kq == 0 && ok[g]) racc[t1[g] - t0] += z;
        }
    }
}

__global__ __launch_bounds__(64 * RHO_WAVES) void k_rho_m64(
    const RhoUnit *__restrict__ units, const int32_t *__restrict__ order, int32_t nunits,
    const Seg *__restrict__ segs, const PairDev *__restrict__ pairs, const double *__restrict__ tab,
    const int32_t *__restrict__ olg1, const int32_t *__restrict__ olg2, const double *__restrict__ Deff,
    const int64_t *__restrict__ gl_off, double *__restrict__ rhoA) {
    __shared__ double racc[RHO_WAVES][RHO_CHUNK];  // the 2048-entry decomposition
    const int b = blockIdx.x;
    if (b >= nunits) return;
    const RhoUnit u = units[order[b]];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int t = threadIdx.x; t < RHO_WAVES * RHO_CHUNK; t += blockDim.x) (&racc[0][0])[t] = 0.0;
    __syncthreads();
    for (int s = u.seg_beg + wave; s < u.seg_end; s += RHO_WAVES) {
        const Seg sg = segs[s];
        const PairDev P = pairs[sg.pair];
        const double *D = Deff + P.hoff;
        const double *tabA = tab + P.tab_a, *tabB = tab + P.tab_b;
        const int32_t *L1 = olg1 + P.olg, *L2 = olg2 + P.olg;
        if (P.nob <= 20) rho_seg_m64<5>(D, tabA, tabB, L1, L2, sg.lo, sg.hi, u.t0, P.noa, P.nob, lane, racc[wave]);
        else rho_seg_m64<8>(D, tabA, tabB, L1, L2, sg.lo, sg.hi, u.t0, P.noa, P.nob, lane, racc[wave]);
    }
    __syncthreads();
    double *out = rhoA + gl_off[u.atom] + u.t0;
    for (int t = threadIdx.x; t < u.nt; t += blockDim.x) {
        double acc = racc[0][t];
#pragma unroll
        for (int w = 1; w < RHO_WAVES; w++) acc += racc[w][t];
        out[t] = acc;
    }
}

// ---- FP64 Force3 on the matrix cores: same Y = D x Phi_B, then the three
// gradient dots weighted by V at the point.
template <int KS>
__device__ __forceinline__ void f3_seg_m64(const double *__restrict__ D, const double *__restrict__ tabB,
                                           const double *__restrict__ dx, const double *__restrict__ dy,
                                           const double *__restrict__ dz, const int32_t *__restrict__ L1,
                                           const int32_t *__restrict__ L2, const int32_t *__restrict__ glA,
                                           const double *__restrict__ V, int n, int noa, int nob, int lane,
                                           double &sx, double &sy, double &sz) {
    const int r16 = lane & 15, kq = lane >> 4;
    double a[2][KS];
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
        for (int st = 0; st < KS; st++) {
            const int i = rt * 16 + r16, j = 4 * st + kq;
            a[rt][st] = (i < noa && j < nob) ? D[i * nob + j] : 0.0;
        }
    for (int base = 0; base < n; base += 16) {
        const int idx = base + r16;
        const bool ok = idx < n;
        const int t1 = L1[ok ? idx : 0], t2 = L2[ok ? idx : 0];
        const double *cb = tabB + (int64_t)t2 * nob;
        double b[KS];
#pragma unroll
        for (int st = 0; st < KS; st++) {
            const int j = 4 * st + kq;
            b[st] = (ok && j < nob) ? cb[j] : 0.0;
        }
        d4_m y[2];
#pragma unroll
        for (int rt = 0; rt < 2; rt++) {
            y[rt] = d4_m{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int st = 0; st < KS; st++) y[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt][st], b[st], y[rt], 0, 0, 0);
        }
        const int64_t ca = (int64_t)t1 * noa;
        double lx = 0.0, ly = 0.0, lz = 0.0;
#pragma unroll
        for (int rt = 0; rt < 2; rt++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = rt * 16 + kq + 4 * v;
                if (i < noa) {
                    lx = fma(dx[ca + i], y[rt][v], lx);
                    ly = fma(dy[ca + i], y[rt][v], ly);
                    lz = fma(dz[ca + i], y[rt][v], lz);
                }
            }
        const double pot = ok ? V[glA[t1]] : 0.0;
        sx = fma(pot, lx, sx);
        sy = fma(pot, ly, sy);
        sz = fma(pot, lz, sz);
    }
}

__global__ __launch_bounds__(64) void k_f3_m64(const F3Unit *__restrict__ units, const int32_t *__restrict__ order,
                                               int32_t nunits, const PairDev *__restrict__ pairs,
                                               const double *__restrict__ tab, const double *__restrict__ dx,
                                               const double *__restrict__ dy, const double *__restrict__ dz,
                                               const int32_t *__restrict__ olg1, const int32_t *__restrict__ olg2,
                                               const int32_t *__restrict__ glist, const double *__restrict__ V,
                                               const double *__restrict__ D, double *__restrict__ fpart) {
    const int b = blockIdx.x;
    if (b >= nunits) return;
    const int ui = order[b];
    if (ui < 0) return;
    const F3Unit u = units[ui];
    const PairDev P = pairs[u.pair];
    const int lane = threadIdx.x;
    const double *Dp = D + P.hoff;
    const int32_t *L1 = olg1 + P.olg + u.lo, *L2 = olg2 + P.olg + u.lo;
    const int32_t *glA = glist + P.gl_a;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    if (P.nob <= 20)
        f3_seg_m64<5>(Dp, tab + P.tab_b, dx + P.tab_a, dy + P.tab_a, dz + P.tab_a, L1, L2, glA, V, u.n, P.noa, P.nob,
                      lane, sx, sy, sz);
    else
        f3_seg_m64<8>(Dp, tab + P.tab_b, dx + P.tab_a, dy + P.tab_a, dz + P.tab_a, L1, L2, glA, V, u.n, P.noa, P.nob,
                      lane, sx, sy, sz);
    for (int off = 32; off >= 1; off >>= 1) {  // fixed butterfly order
        sx += __shfl_xor(sx, off, 64);
        sy += __shfl_xor(sy, off, 64);
        sz += __shfl_xor(sz, off, 64);
    }
    if (lane == 0) {
        fpart[3 * (int64_t)ui] = sx;
        fpart[3 * (int64_t)ui + 1] = sy;
        fpart[3 * (int64_t)ui + 2] = sz;
    }
}

// rho(g) = sum of the per-entry partials of the entries at g (CSR order).
__global__ void k_rho_gather(const int64_t *__restrict__ ent_off, const int64_t *__restrict__ ent,
                             const double *__restrict__ rhoA, int64_t nn, double *__restrict__ rho) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nn) return;
    double s = 0.0;
    for (int64_t e = ent_off[g]; e < ent_off[g + 1]; e++) s += rhoA[ent[e]];
    rho[g] = s;
}

// ---- Force3: point work of one unit (lane = overlap point) --------------------
template <typename T, int NA, int NB>
__device__ __forceinline__ void f3_points(const T *__restrict__ D, const T *__restrict__ tabA,
                                          const T *__restrict__ tabB, const T *__restrict__ dx,
                                          const T *__restrict__ dy, const T *__restrict__ dz,
                                          const int32_t *__restrict__ L1, const int32_t *__restrict__ L2,
                                          const int32_t *__restrict__ glA, const double *__restrict__ V, int n,
                                          int lane, double &sx, double &sy, double &sz) {
    for (int base = 0; base < n; base += 64) {
        const int idx = base + lane;
        const bool ok = idx < n;
        const int t1 = L1[ok ? idx : 0], t2 = L2[ok ? idx : 0];
        const T *cb = tabB + (int64_t)t2 * NB;
        const int64_t ca = (int64_t)t1 * NA;
        T bv[NB];
#pragma unroll
        for (int j = 0; j < NB; j++) bv[j] = cb[j];
        T lx = 0, ly = 0, lz = 0;
#pragma unroll
        for (int i = 0; i < NA; i++) {
            T y = 0;
#pragma unroll
            for (int j = 0; j < NB; j++) y = fma(D[i * NB + j], bv[j], y);
            lx = fma(dx[ca + i], y, lx);
            ly = fma(dy[ca + i], y, ly);
            lz = fma(dz[ca + i], y, lz);
        }
        const double pot = ok ? V[glA[t1]] : 0.0;
        sx = fma(pot, (double)lx, sx);
        sy = fma(pot, (double)ly, sy);
        sz = fma(pot, (double)lz, sz);
    }
}

template <typename T>
__device__ void f3_points_any(const T *__restrict__ D, const T *__restrict__ tabA, const T *__restrict__ tabB,
                              const T *__restrict__ dx, const T *__restrict__ dy, const T *__restrict__ dz,
                              const int32_t *__restrict__ L1, const int32_t *__restrict__ L2,
                              const int32_t *__restrict__ glA, const double *__restrict__ V, int n, int lane,
                              double &sx, double &sy, double &sz, int noa, int nob) {
    for (int base = 0; base < n; base += 64) {
        const int idx = base + lane;
        const bool ok = idx < n;
        const int t1 = L1[ok ? idx : 0], t2 = L2[ok ? idx : 0];
        const T *cb = tabB + (int64_t)t2 * nob;
        const int64_t ca = (int64_t)t1 * noa;
        T lx = 0, ly = 0, lz = 0;
        for (int i = 0; i < noa; i++) {
            T y = 0;
            for (int j = 0; j < nob; j++) y = fma(D[i * nob + j], cb[j], y);
            lx = fma(dx[ca + i], y, lx);
            ly = fma(dy[ca + i], y, ly);
            lz = fma(dz[ca + i], y, lz);
        }
        const double pot = ok ? V[glA[t1]] : 0.0;
        sx = fma(pot, (double)lx, sx);
        sy = fma(pot, (double)ly, sy);
        sz = fma(pot, (double)lz, sz);
    }
}

template <typename T>
__global__ __launch_bounds__(64) void k_f3(const F3Unit *__restrict__ units, const int32_t *__restrict__ order,
                                           int32_t nunits, const PairDev *__restrict__ pairs,
                                           const T *__restrict__ tab, const T *__restrict__ dx,
                                           const T *__restrict__ dy, const T *__restrict__ dz,
                                           const int32_t *__restrict__ olg1, const int32_t *__restrict__ olg2,
                                           const int32_t *__restrict__ glist, const double *__restrict__ V,
                                           const T *__restrict__ D, double *__restrict__ fpart) {
    const int b = blockIdx.x;
    if (b >= nunits) return;
    const int ui = order[b];
    if (ui < 0) return;
    const F3Unit u = units[ui];
    const PairDev P = pairs[__builtin_amdgcn_readfirstlane(u.pair)];
    const int noa = __builtin_amdgcn_readfirstlane(P.noa), nob = __builtin_amdgcn_readfirstlane(P.nob);
    const int lane = threadIdx.x;
    const T *Dp = D + P.hoff;
    const int32_t *L1 = olg1 + P.olg + u.lo, *L2 = olg2 + P.olg + u.lo;
    const int32_t *glA = glist + P.gl_a;
    const T *tA = tab + P.tab_a, *tB = tab + P.tab_b;
    const T *ax = dx + P.tab_a, *ay = dy + P.tab_a, *az = dz + P.tab_a;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    bool done = false;
#define F3_CASE(A, B)                                                                              \
    if (!done && noa == A && nob == B) {                                                           \
        f3_points<T, A, B>(Dp, tA, tB, ax, ay, az, L1, L2, glA, V, u.n, lane, sx, sy, sz);         \
        done = true;                                                                               \
    }
    NO_CASES(F3_CASE)
#undef F3_CASE
    if (!done) f3_points_any<T>(Dp, tA, tB, ax, ay, az, L1, L2, glA, V, u.n, lane, sx, sy, sz, noa, nob);
    for (int off = 32; off >= 1; off >>= 1) {  // fixed butterfly order
        sx += __shfl_xor(sx, off, 64);
        sy += __shfl_xor(sy, off, 64);
        sz += __shfl_xor(sz, off, 64);
    }
    if (lane == 0) {
        fpart[3 * (int64_t)ui] = sx;
        fpart[3 * (int64_t)ui + 1] = sy;
        fpart[3 * (int64_t)ui + 2] = sz;
    }
}

// Per-atom sums of Force3 unit partials (pairs are atom-major, units
// consecutive per pair): fsum[a][xyz] += w * sum.  One wave per atom: lane l
// adds units l, l + 64, ... in order, then a fixed butterfly (deterministic).
__global__ __launch_bounds__(256) void k_f3_reduce(const int32_t *__restrict__ atom_ub, int natom,
                                                   const double *__restrict__ fpart, double w,
                                                   double *__restrict__ fsum) {
    const int a = blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (a >= natom) return;  // whole waves
    double s[3] = {0.0, 0.0, 0.0};
    for (int u = atom_ub[a] + lane; u < atom_ub[a + 1]; u += 64)
        for (int x = 0; x < 3; x++) s[x] += fpart[3 * (int64_t)u + x];
    for (int off = 32; off >= 1; off >>= 1)
        for (int x = 0; x < 3; x++) s[x] += __shfl_xor(s[x], off, 64);
    if (lane == 0)
        for (int x = 0; x < 3; x++) fsum[3 * a + x] += w * s[x];
}

// force[a + natom*xyz] += fsum[a][xyz] * GridVol
__global__ void k_f3_apply(int natom, const double *__restrict__ fsum, double gridvol, double *__restrict__ force) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natom) return;
    for (int x = 0; x < 3; x++) force[a + (int64_t)natom * x] += fsum[3 * a + x] * gridvol;
}

// ---------------------------------------------------------------------------
// Table generation: Set_Orbitals_Grid! (Set_Orbitals_Grid.jl:45-277) and
// Set_dOrbitals_Grid! (Set_dOrbitals_Grid.jl:23-362), one thread per grid
// list entry.  Expression order follows the reference; compiled without FP
// contraction so the tables match the host restatement to the last bits of
// the math library.
// ---------------------------------------------------------------------------
struct TabGeom {
    double glat[9], origin[3];
    int32_t n2, n3, natom;
};

__device__ inline void xyz_to_spherical_d(double x, double y, double z, double &r, double &theta, double &phi) {
    const double Min_r = 1e-14;
    const double pi = 3.141592653589793;
    const double dum = x * x + y * y;
    r = sqrt(dum + z * z);
    const double r1 = sqrt(dum);
    if (Min_r <= r) {
        const double sz = z > 0.0 ? 1.0 : (z < 0.0 ? -1.0 : z);
        theta = acos(r < fabs(z) ? sz : z / r);
        if (Min_r <= r1) {
            const double sy = y > 0.0 ? 1.0 : (y < 0.0 ? -1.0 : y);
            const double d2 = r1 < fabs(y) ? sy : y / r1;
            phi = x >= 0.0 ? asin(d2) : pi - asin(d2);
        } else {
            phi = 0.0;
        }
    } else {
        theta = 0.5 * pi;
        phi = 0.0;
    }
}

// Real spherical harmonics and their angular derivatives in the reference's
// m order per l (Set_Orbitals_Grid.jl:243-267; Ylm_real.jl:1-162).
__device__ inline void ylm_all(int l, double siQ, double coQ, double siP, double coP, double *Y, double *dQ,
                               double *dP, bool deriv) {
    const double pi = 3.141592653589793;
    if (l == 0) {
        Y[0] = 0.5 * sqrt(1.0 / pi);
        if (deriv) { dQ[0] = 0.0; dP[0] = 0.0; }
    } else if (l == 1) {
        const double c = sqrt(3.0 / 4.0 / pi);
        Y[0] = c * siQ * coP; Y[1] = c * siQ * siP; Y[2] = c * coQ;   // px, py, pz
        if (deriv) {
            dQ[0] = c * coQ * coP; dQ[1] = c * coQ * siP; dQ[2] = -c * siQ;
            dP[0] = -c * siP; dP[1] = c * coP; dP[2] = 0.0;
        }
    } else if (l == 2) {
        const double c15 = sqrt(15.0 / pi), c5 = sqrt(5.0 / pi);
        // dz2, dx2-y2, dxy, dxz, dyz
        Y[0] = 0.25 * c5 * (3 * (coQ * coQ) - 1);
        Y[1] = 0.25 * c15 * (siQ * siQ) * (1 - 2 * (siP * siP));
        Y[2] = 0.5 * c15 * (siQ * siQ) * siP * coP;
        Y[3] = 0.5 * c15 * siQ * coQ * coP;
        Y[4] = 0.5 * c15 * siQ * coQ * siP;
        if (deriv) {
            dQ[0] = -1.5 * c5 * siQ * coQ;
            dQ[1] = 0.5 * c15 * siQ * coQ * (1 - 2 * (siP * siP));
            dQ[2] = c15 * siQ * coQ * siP * coP;
            dQ[3] = 0.5 * c15 * (1 - 2 * (siQ * siQ)) * coP;
            dQ[4] = 0.5 * c15 * (1 - 2 * (siQ * siQ)) * siP;
            dP[0] = 0.0;
            dP[1] = -c15 * siQ * siP * coP;
            dP[2] = 0.5 * c15 * siQ * (1 - 2 * (siP * siP));
            dP[3] = -0.5 * c15 * coQ * siP;
            dP[4] = 0.5 * c15 * coQ * coP;
        }
    } else {
        const double c7 = sqrt(7.0 / pi), c21 = sqrt(21.0 / 2.0 / pi), c105 = sqrt(105.0 / pi), c35 = sqrt(35.0 / 2.0 / pi);
        // z3, xz2, yz2, z(x2-y2), xyz, x(x2-3y2), y(3x2-y2)
        Y[0] = 0.25 * c7 * (5 * (coQ * coQ * coQ) - 3 * coQ);
        Y[1] = 0.25 * c21 * siQ * (5 * (coQ * coQ) - 1) * coP;
        Y[2] = 0.25 * c21 * siQ * (5 * (coQ * coQ) - 1) * siP;
        Y[3] = 0.25 * c105 * (siQ * siQ) * coQ * ((coP * coP) - (siP * siP));
        Y[4] = 0.5 * c105 * (siQ * siQ) * coQ * siP * coP;
        Y[5] = 0.25 * c35 * (siQ * siQ * siQ) * (4 * (coP * coP * coP) - 3 * coP);
        Y[6] = 0.25 * c35 * (siQ * siQ * siQ) * (3 * siP - 4 * (siP * siP * siP));
        if (deriv) {
            dQ[0] = 0.25 * c7 * siQ * (-15 * (coQ * coQ) + 3);
            dQ[1] = 0.25 * c21 * coP * coQ * (15 * (coQ * coQ) - 11);
            dQ[2] = 0.25 * c21 * siP * coQ * (15 * (coQ * coQ) - 11);
            dQ[3] = 0.25 * c105 * ((coP * coP) - (siP * siP)) * siQ * (2 * (coQ * coQ) - (siQ * siQ));
            dQ[4] = 0.5 * c105 * coP * siP * siQ * (2 * (coQ * coQ) - (siQ * siQ));
            dQ[5] = 0.75 * c35 * coP * coQ * (siQ * siQ) * (-3 + 4 * (coP * coP));
            dQ[6] = 0.75 * c35 * (siQ * siQ) * coQ * siP * (3 - 4 * (siP * siP));
            dP[0] = 0.0;
            dP[1] = 0.25 * c21 * siP * (-5 * (coQ * coQ) + 1);
            dP[2] = 0.25 * c21 * coP * (5 * (coQ * coQ) - 1);
            dP[3] = -c105 * coP * coQ * siP * siQ;
            dP[4] = 0.5 * c105 * coQ * siQ * ((coP * coP) - (siP * siP));
            dP[5] = 0.75 * c35 * siP * (siQ * siQ) * (1 - 4 * (coP * coP));
            dP[6] = 0.75 * c35 * coP * (siQ * siQ) * (1 - 4 * (siP * siP));
        }
    }
}

// The Ylm constants are formed as products with sqrt() of the literal
// quotients, e.g. 0.5*sqrt(15/pi)*siQ*coQ*coP = ((0.5*c15)*siQ)*coQ*coP,
// which is how the expressions above associate.
__global__ void k_tables(int64_t nent, const int64_t *__restrict__ gl_off, const int32_t *__restrict__ glist,
                         const int32_t *__restrict__ clist, const int64_t *__restrict__ tab_off,
                         const double *__restrict__ gxyz, const int32_t *__restrict__ a2s,
                         const double *__restrict__ atv, TabGeom tg, const SpeciesDev *__restrict__ spe,
                         int deriv, double *__restrict__ tab, double *__restrict__ tx, double *__restrict__ ty,
                         double *__restrict__ tz, int32_t *__restrict__ flag) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nent) return;
    // owning atom by binary search over gl_off
    int lo = 0, hi = tg.natom;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (gl_off[mid] <= e) lo = mid; else hi = mid;
    }
    const int a = lo;
    const int t = (int)(e - gl_off[a]);
    const SpeciesDev &S = spe[a2s[a]];
    const int GNc = glist[e], GRc = clist[e];
    const int n1 = GNc / (tg.n2 * tg.n3);
    const int n2 = (GNc - n1 * tg.n2 * tg.n3) / tg.n3;
    const int n3 = GNc - n1 * tg.n2 * tg.n3 - n2 * tg.n3;
    const double Cx = n1 * tg.glat[0] + n2 * tg.glat[3] + n3 * tg.glat[6] + tg.origin[0];
    const double Cy = n1 * tg.glat[1] + n2 * tg.glat[4] + n3 * tg.glat[7] + tg.origin[1];
    const double Cz = n1 * tg.glat[2] + n2 * tg.glat[5] + n3 * tg.glat[8] + tg.origin[2];
    double x = Cx + atv[3 * GRc] - gxyz[3 * a];
    double y = Cy + atv[3 * GRc + 1] - gxyz[3 * a + 1];
    double z = Cz + atv[3 * GRc + 2] - gxyz[3 * a + 2];
    double R, theta, phi;
    xyz_to_spherical_d(x, y, z, R, theta, phi);
    const double Rmin = 1.0e-13;
    if (deriv && R < Rmin) {  // Set_dOrbitals_Grid.jl:107-112
        x = x + Rmin; y = y + Rmin; z = z + Rmin;
        xyz_to_spherical_d(x, y, z, R, theta, phi);
    }
    const double *RV = S.rv;
    const int nm = S.nmesh;
    double RF[4][16], dRF[4][16];
    bool outside = false;
    if (RV[nm - 1] < R) {
        outside = true;
        for (int l = 0; l <= S.maxl; l++)
            for (int p = 0; p < S.num_basis[l]; p++) { RF[l][p] = 0.0; dRF[l][p] = 0.0; }
    } else if (R < RV[0]) {
        // polynomial continuation inside the first mesh point (Set_Orbitals_Grid.jl:124-185)
        const int m = 4;
        const double rm = RV[m];
        const double h1 = RV[m - 1] - RV[m - 2], h2 = RV[m] - RV[m - 1], h3 = RV[m + 1] - RV[m];
        const double x1 = rm - RV[m - 1], x2 = rm - RV[m];
        const double y1 = x1 / h2, y2 = x2 / h2, y12 = y1 * y1, y22 = y2 * y2;
        double dum = h1 + h2;
        const double dum1 = h1 / h2 / dum, dum2 = h2 / h1 / dum;
        dum = h2 + h3;
        const double dum3 = h2 / h3 / dum, dum4 = h3 / h2 / dum;
        for (int l = 0; l <= S.maxl; l++)
            for (int p = 0; p < S.num_basis[l]; p++) {
                const double *W = S.rwf[l][p];
                const double f1 = W[m - 2], f2 = W[m - 1], f3 = W[m], f4 = W[m + 1];
                const double d = f3 - f2;
                const double g1 = d * dum1 + (f2 - f1) * dum2;
                const double g2 = (f4 - f3) * dum3 + d * dum4;
                const double A = 3 * f2 + h2 * g1, B = 2 * f2 + h2 * g1;
                const double Cc = 3 * f3 - h2 * g2, D = 2 * f3 - h2 * g2;
                const double E = A + B * y2, F = Cc - D * y1;
                const double f = y22 * E + y12 * F;
                const double df = deriv ? (2 * y2 * E + y22 * B + 2.0 * y1 * F - y12 * D) / h2
                                        : (2 * y2 / h2) * E + y22 * B / h2 + (2 * y1 / h2) * F - y12 * D / h2;
                double ca, cb, cc, cd;
                if (l == 0) { ca = 0.0; cb = 0.5 * df / rm; cc = 0.0; cd = f - cb * (rm * rm); }
                else if (l == 1) { ca = (rm * df - f) / (2 * (rm * rm * rm)); cb = 0.0; cc = df - 3 * ca * (rm * rm); cd = 0.0; }
                else { cb = (3 * f - rm * df) / (rm * rm); ca = (f - cb * (rm * rm)) / (rm * rm * rm); cc = 0.0; cd = 0.0; }
                RF[l][p] = ca * (R * R * R) + cb * (R * R) + cc * R + cd;
                dRF[l][p] = 3 * ca * (R * R) + 2 * cb * R + cc;
            }
    } else {
        int mp_min = 1, mp_max = nm;  // 1-based, as in the reference's search
        while (mp_max - mp_min != 1) {
            const int mm = (mp_min + mp_max) / 2;
            if (RV[mm - 1] < R) mp_min = mm; else mp_max = mm;
        }
        const int m = mp_max;
        if (m < 3 || m + 1 > nm) {  // the reference would read past the mesh here
            atomicOr(flag, 1);
            return;
        }
        const double h1 = RV[m - 2] - RV[m - 3], h2 = RV[m - 1] - RV[m - 2], h3 = RV[m] - RV[m - 1];
        const double x1 = R - RV[m - 2], x2 = R - RV[m - 1];
        const double y1 = x1 / h2, y2 = x2 / h2, y12 = y1 * y1, y22 = y2 * y2;
        double dum = h1 + h2;
        const double dum1 = h1 / h2 / dum, dum2 = h2 / h1 / dum;
        dum = h2 + h3;
        const double dum3 = h2 / h3 / dum, dum4 = h3 / h2 / dum;
        for (int l = 0; l <= S.maxl; l++)
            for (int p = 0; p < S.num_basis[l]; p++) {
                const double *W = S.rwf[l][p];
                double f1 = W[m - 3], f4 = W[m];
                const double f2 = W[m - 2], f3 = W[m - 1];
                if (!deriv && m == nm - 1) f4 = f1;       // Set_Orbitals_Grid.jl:226-228
                if (deriv && m == nm) f4 = f1;            // Set_dOrbitals_Grid.jl:230-232
                const double d = f3 - f2;
                const double g1 = d * dum1 + (f2 - f1) * dum2;
                const double g2 = (f4 - f3) * dum3 + d * dum4;
                if (!deriv) {
                    RF[l][p] = (y22 * (3 * f2 + h2 * g1 + (2 * f2 + h2 * g1) * y2) +
                                y12 * (3 * f3 - h2 * g2 - (2 * f3 - h2 * g2) * y1));
                } else {
                    const double A = 3 * f2 + h2 * g1, B = 2 * f2 + h2 * g1;
                    const double Cc = 3 * f3 - h2 * g2, D = 2 * f3 - h2 * g2;
                    const double E = A + B * y2, F = Cc - D * y1;
                    RF[l][p] = y22 * E + y12 * F;
                    dRF[l][p] = (2 * y2 * E + y22 * B + 2.0 * y1 * F - y12 * D) / h2;
                }
            }
    }
    const int NO = S.no;
    const int64_t base = tab_off[a] + (int64_t)t * NO;
    double siQ = 0.0, coQ = 0.0, siP = 0.0, coP = 0.0;
    if (!outside) { siQ = sin(theta); coQ = cos(theta); siP = sin(phi); coP = cos(phi); }
    double Y[7], dQ[7], dP[7];
    int ist = 0;
    if (!deriv) {
        for (int l = 0; l <= S.maxl; l++) {
            ylm_all(l, siQ, coQ, siP, coP, Y, dQ, dP, false);
            for (int p = 0; p < S.num_basis[l]; p++)
                for (int mi = 0; mi < 2 * l + 1; mi++) tab[base + ist++] = outside ? 0.0 : RF[l][p] * Y[mi];
        }
        return;
    }
    double dRx = 0, dRy = 0, dRz = 0, dQx = 0, dQy = 0, dQz = 0, dPx = 0, dPy = 0, dPz = 0;
    if (!outside) {
        dRx = siQ * coP; dRy = siQ * siP; dRz = coQ;
        if (R > Rmin) {
            dQx = coQ * coP / R; dQy = coQ * siP / R; dQz = -siQ / R;
            dPx = -siP / R; dPy = coP / R; dPz = 0.0;
        }
    }
    for (int l = 0; l <= S.maxl; l++) {
        ylm_all(l, siQ, coQ, siP, coP, Y, dQ, dP, true);
        for (int p = 0; p < S.num_basis[l]; p++)
            for (int mi = 0; mi < 2 * l + 1; mi++) {
                if (outside) { tx[base + ist] = 0.0; ty[base + ist] = 0.0; tz[base + ist] = 0.0; ist++; continue; }
                const double dchidr = dRF[l][p] * Y[mi];
                const double dchidtheta = RF[l][p] * dQ[mi];
                const double dchidphi = RF[l][p] * dP[mi];
                tx[base + ist] = -dRx * dchidr - dQx * dchidtheta - dPx * dchidphi;
                ty[base + ist] = -dRy * dchidr - dQy * dchidtheta - dPy * dchidphi;
                tz[base + ist] = -dRz * dchidr - dQz * dchidtheta - dPz * dchidphi;
                ist++;
            }
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
namespace {

template <class T>
int upload(lcpao_ctx *ctx, T **dst, const std::vector<T> &src) {
    if (*dst) { (void)hipFree(*dst); *dst = nullptr; }
    const size_t bytes = std::max<size_t>(1, src.size()) * sizeof(T);
    HIP_TRY(ctx, hipMalloc(reinterpret_cast<void **>(dst), bytes));
    if (!src.empty()) HIP_TRY(ctx, hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return LCPAO_OK;
}

int dev_alloc(lcpao_ctx *ctx, void **p, size_t bytes) {
    if (*p) { (void)hipFree(*p); *p = nullptr; }
    HIP_TRY(ctx, hipMalloc(p, std::max<size_t>(1, bytes)));
    return LCPAO_OK;
}

// cell index -> (i,j,k) for the Generation_ATV! order (cell 0 = home, then
// the (i,j,k) loop over [-C,C]^3 skipping (0,0,0))
void cell_ijk(int cpy, int cell, int *ijk) {
    if (cell == 0) { ijk[0] = ijk[1] = ijk[2] = 0; return; }
    const int W = 2 * cpy + 1;
    const int zero_lin = (cpy * W + cpy) * W + cpy;
    const int lin = cell <= zero_lin ? cell - 1 : cell;
    ijk[0] = lin / (W * W) - cpy;
    ijk[1] = (lin / W) % W - cpy;
    ijk[2] = lin % W - cpy;
}

int cell_index(int cpy, const int *ijk) {
    if (ijk[0] == 0 && ijk[1] == 0 && ijk[2] == 0) return 0;
    const int W = 2 * cpy + 1;
    const int lin = ((ijk[0] + cpy) * W + (ijk[1] + cpy)) * W + (ijk[2] + cpy);
    const int zero_lin = (cpy * W + cpy) * W + cpy;
    return lin < zero_lin ? lin + 1 : lin;
}

int set_err(lcpao_ctx *ctx, int code, const std::string &m) {
    ctx->err = m;
    return code;
}


// Orders unit indices by the grid index of their first point.
template <class GetKey>
std::vector<int32_t> launch_order(size_t n, GetKey key) {
    std::vector<int64_t> k(n);
    for (size_t i = 0; i < n; i++) k[i] = key(i);
    std::vector<int32_t> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int32_t x, int32_t y) { return k[x] < k[y]; });
    return ord;
}

constexpr int NXCD = 8;  // XCDs of an MI355X (workgroup w runs on XCD w % 8)

// Deals units in launch order to the eight XCDs in contiguous ranges:
// workgroup w runs on XCD w % 8, so entry 8 j + k is unit j of range k and
// each XCD walks its own band of grid planes, whose table columns the pairs
// crossing it gather again from that XCD's L2 (-1: no unit, for ranges of
// unequal length).
std::vector<int32_t> xcd_deal(const std::vector<int32_t> &sorted) {
    const size_t n = sorted.size(), per = (n + NXCD - 1) / NXCD;
    std::vector<int32_t> out(per * NXCD, -1);
    for (size_t k = 0; k < (size_t)NXCD; k++)
        for (size_t j = 0; j < per && k * per + j < n; j++) out[j * NXCD + k] = sorted[k * per + j];
    return out;
}

int build_work(lcpao_ctx *ctx, const lcpao_geom *g, const std::vector<int32_t> &o1, const std::vector<int32_t> &o2,
               const std::vector<int32_t> &glist_h) {
    const int64_t np = ctx->npair;
    // transpose partner of each local pair: (B, A, -Rn)
    std::vector<std::vector<std::pair<int64_t, int>>> by_atom(ctx->natom);
    auto key = [](int b, int cell) { return (int64_t)b * 4000037LL + cell; };
    for (int64_t p = 0; p < np; p++) by_atom[ctx->pairs[p].a].push_back({key(ctx->pairs[p].b, g->pair_ncn[p]), (int)p});
    for (auto &v : by_atom) std::sort(v.begin(), v.end());
    for (int64_t p = 0; p < np; p++) {
        int ijk[3], nijk[3];
        cell_ijk(ctx->cpy, g->pair_ncn[p], ijk);
        for (int d = 0; d < 3; d++) nijk[d] = -ijk[d];
        const int64_t k = key(ctx->pairs[p].a, cell_index(ctx->cpy, nijk));
        const auto &v = by_atom[ctx->pairs[p].b];
        auto it = std::lower_bound(v.begin(), v.end(), std::make_pair(k, -1));
        ctx->pairs[p].partner = (it != v.end() && it->first == k) ? it->second : -1;
    }
    // computed pairs: partner absent or not smaller (the transpose is then filled)
    std::vector<int32_t> &canon = ctx->canon;
    canon.clear();
    for (int64_t p = 0; p < np; p++) {
        const int q = ctx->pairs[p].partner;
        if (q < 0 || q >= p) canon.push_back((int32_t)p);
    }
    ctx->ncanon = (int64_t)canon.size();
    auto first_g = [&](int p, int lo) -> int64_t {
        const PairDev &P = ctx->pairs[p];
        if (g->numolg[p] == 0) return INT64_MAX;
        return glist_h[P.gl_a + o1[P.olg + lo]];
    };
    // H units over computed pairs, contiguous per pair
    ctx->hunits.clear();
    std::vector<int32_t> hub(1, 0);
    int64_t part = 0;
    int max_ne = 1;
    for (int p : canon) {
        const PairDev &P = ctx->pairs[p];
        const int n = g->numolg[p];
        for (int lo = 0; lo < n; lo += HUNIT_PTS) {
            ctx->hunits.push_back({p, lo, std::min(HUNIT_PTS, n - lo), 0, part});
            part += (int64_t)P.noa * P.nob;
        }
        max_ne = std::max(max_ne, P.noa * P.nob);
        hub.push_back((int32_t)ctx->hunits.size());
    }
    ctx->hred_threads = std::min(1024, (max_ne + 63) / 64 * 64);
    ctx->hpart_size = part;
    // (in contiguous XCD ranges: GaAs216 FP64 H 4.69 against 5.81 ms)
    auto horder = xcd_deal(launch_order(ctx->hunits.size(), [&](size_t i) { return first_g(ctx->hunits[i].pair, ctx->hunits[i].lo); }));
    ctx->horder_n = (int32_t)horder.size();
    // Force3 units over all pairs; pairs are grouped by home atom
    ctx->funits.clear();
    std::vector<int32_t> aub(ctx->natom + 1, 0);
    for (int64_t p = 0; p < np; p++) {
        if (p > 0 && ctx->pairs[p].a < ctx->pairs[p - 1].a) return set_err(ctx, LCPAO_EINVAL, "pairs must be grouped by home atom in ascending order");
        const int n = g->numolg[p];
        for (int lo = 0; lo < n; lo += UNIT_PTS) {
            ctx->funits.push_back({(int32_t)p, lo, std::min(UNIT_PTS, n - lo), 0});
            aub[ctx->pairs[p].a + 1]++;
        }
    }
    for (int a = 0; a < ctx->natom; a++) aub[a + 1] += aub[a];
    // (in contiguous XCD ranges: Si511 FP64 Force3 41.7 against 44.6 ms)
    auto forder = xcd_deal(launch_order(ctx->funits.size(), [&](size_t i) { return first_g(ctx->funits[i].pair, ctx->funits[i].lo); }));
    ctx->forder_n = (int32_t)forder.size();
    // density units of each decomposition: (atom, chunk) holding the
    // segments of its computed pairs
    std::vector<std::vector<int32_t>> apairs(ctx->natom);
    for (int p : canon) apairs[ctx->pairs[p].a].push_back(p);
    int rc;
    for (int d = 0; d < RHO_NDECOMP; d++) {
        const int CH = RHO_DECOMP_CHUNK[d];
        std::vector<RhoUnit> runits;
        std::vector<Seg> segs;
        for (int a = 0; a < ctx->natom; a++) {
            std::vector<int> cur(apairs[a].size(), 0);
            for (int t0 = 0; t0 < ctx->gridn[a]; t0 += CH) {
                RhoUnit u;
                u.atom = a; u.t0 = t0; u.nt = std::min(CH, ctx->gridn[a] - t0);
                u.seg_beg = (int32_t)segs.size(); u.pad = 0;
                for (size_t k = 0; k < apairs[a].size(); k++) {
                    const int p = apairs[a][k];
                    const int32_t *L1 = o1.data() + ctx->pairs[p].olg;
                    const int n = g->numolg[p];
                    const int lo = cur[k];
                    int hi = lo;
                    while (hi < n && L1[hi] < t0 + u.nt) hi++;
                    if (hi > lo) segs.push_back({p, lo, hi, 0});
                    cur[k] = hi;
                }
                u.seg_end = (int32_t)segs.size();
                runits.push_back(u);
            }
        }
        if (segs.size() > (size_t)INT32_MAX) return set_err(ctx, LCPAO_ERANGE, "too many density segments");
        auto rorder = launch_order(runits.size(), [&](size_t i) { return (int64_t)glist_h[ctx->gl_off[runits[i].atom] + runits[i].t0]; });
        auto &D = ctx->rdec[d];
        D.nunits = (int64_t)runits.size();
        if ((rc = upload(ctx, &D.d_units, runits))) return rc;
        if ((rc = upload(ctx, &D.d_order, rorder))) return rc;
        if ((rc = upload(ctx, &D.d_segs, segs))) return rc;
    }
    if ((rc = upload(ctx, &ctx->d_hunits, ctx->hunits))) return rc;
    if ((rc = upload(ctx, &ctx->d_horder, horder))) return rc;
    if ((rc = upload(ctx, &ctx->d_hcanon, canon))) return rc;
    if ((rc = upload(ctx, &ctx->d_hpair_ub, hub))) return rc;
    if ((rc = upload(ctx, &ctx->d_funits, ctx->funits))) return rc;
    if ((rc = upload(ctx, &ctx->d_forder, forder))) return rc;
    if ((rc = upload(ctx, &ctx->d_atom_ub, aub))) return rc;
    if ((rc = dev_alloc(ctx, (void **)&ctx->d_hpart, (size_t)std::max<int64_t>(1, part) * sizeof(double)))) return rc;
    if ((rc = dev_alloc(ctx, (void **)&ctx->d_fpart, (size_t)std::max<size_t>(1, ctx->funits.size()) * 3 * sizeof(double)))) return rc;
    if ((rc = dev_alloc(ctx, (void **)&ctx->d_fsum, (size_t)ctx->natom * 3 * sizeof(double)))) return rc;
    if ((rc = dev_alloc(ctx, (void **)&ctx->d_deff, (size_t)std::max<int64_t>(1, ctx->my_hsize) * sizeof(double)))) return rc;
    if ((rc = dev_alloc(ctx, (void **)&ctx->d_rhoA, (size_t)std::max<int64_t>(1, ctx->total_gridn) * sizeof(double)))) return rc;
    // entries of each grid point (CSR by grid index, atoms ascending)
    std::vector<int64_t> eoff(ctx->nn + 1, 0), ent(ctx->total_gridn);
    for (int64_t e = 0; e < ctx->total_gridn; e++) eoff[glist_h[e] + 1]++;
    for (int64_t q = 0; q < ctx->nn; q++) eoff[q + 1] += eoff[q];
    {
        std::vector<int64_t> fill(eoff.begin(), eoff.end() - 1);
        for (int64_t e = 0; e < ctx->total_gridn; e++) ent[fill[glist_h[e]]++] = e;
    }
    if ((rc = upload(ctx, &ctx->d_ent_off, eoff))) return rc;
    if ((rc = upload(ctx, &ctx->d_ent, ent))) return rc;
    return LCPAO_OK;
}

int check_ctx(lcpao_ctx *ctx) {
    if (!ctx) return LCPAO_EINVAL;
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, LCPAO_EDEVICE, "hipSetDevice failed");
    return LCPAO_OK;
}

int stage(lcpao_ctx *ctx, int64_t ndoubles) {
    if (ndoubles <= ctx->stage_size) return LCPAO_OK;
    int rc = dev_alloc(ctx, (void **)&ctx->d_stage, (size_t)ndoubles * sizeof(double));
    if (rc) return rc;
    ctx->stage_size = ndoubles;
    return LCPAO_OK;
}

int check_precision(lcpao_ctx *ctx, int precision) {
    if (precision != LCPAO_FP64 && precision != LCPAO_FP32 && precision != LCPAO_BF16X3)
        return set_err(ctx, LCPAO_EINVAL, "precision must be LCPAO_FP64, LCPAO_FP32 or LCPAO_BF16X3");
    return LCPAO_OK;
}


int ham_dev(lcpao_ctx *ctx, int nch, const double *const *d_V, double *const *d_H, int precision) {
    if (!ctx->tables) return set_err(ctx, LCPAO_ESTATE, "orbital tables are not set");
    if (nch != 1 && nch != 2 && nch != 4) return set_err(ctx, LCPAO_EINVAL, "nch must be 1, 2 or 4");
    int rc = check_precision(ctx, precision);
    if (rc) return rc;
    if (precision != LCPAO_FP64) {
        rc = split_hamiltonian(ctx, nch, d_V, d_H, precision);
        if (rc != LCPAO_ERANGE || ctx->block_state != 0) return rc;
        // the block plan cannot hold this geometry: the FP64 pair-list
        // kernels serve the FP32 and bf16x3 H (within any form's bound)
        ctx->err.clear();
    }
    const int nu = (int)ctx->hunits.size();
    for (int c = 0; c < nch; c++) {
        if (!d_V[c] || !d_H[c]) return set_err(ctx, LCPAO_EINVAL, "null channel pointer");
        if (nu > 0) {
            KTimer kt(ctx, KID_HAM);
            if (ctx->maxno <= M4_NO)
                hipLaunchKernelGGL(k_ham_m4, dim3(ctx->horder_n), dim3(64), 0, ctx->stream, ctx->d_hunits, ctx->d_horder,
                                   ctx->horder_n,
                                   ctx->d_pairs, ctx->d_tab, ctx->d_olg1, ctx->d_olg2, ctx->d_glist, d_V[c],
                                   ctx->d_hpart);
            else if (ctx->maxno <= 32)
                hipLaunchKernelGGL(k_ham_mfma64<false>, dim3(ctx->horder_n), dim3(64), 0, ctx->stream, ctx->d_hunits,
                                   ctx->d_horder, ctx->horder_n, ctx->d_pairs, ctx->d_tab, ctx->d_olg1, ctx->d_olg2, ctx->d_glist, d_V[c],
                                   ctx->d_hpart);
            else
                hipLaunchKernelGGL(k_ham_mfma64<true>, dim3(ctx->horder_n), dim3(64), 0, ctx->stream, ctx->d_hunits,
                                   ctx->d_horder, ctx->horder_n, ctx->d_pairs, ctx->d_tab, ctx->d_olg1, ctx->d_olg2, ctx->d_glist, d_V[c],
                                   ctx->d_hpart);
        }
        if (ctx->ncanon > 0) {
            KTimer kt(ctx, KID_HRED);
            hipLaunchKernelGGL(k_ham_reduce, dim3((unsigned)ctx->ncanon), dim3(ctx->hred_threads), 0, ctx->stream, ctx->d_hcanon,
                               ctx->d_hpair_ub, ctx->ncanon, ctx->d_hunits, ctx->d_pairs, ctx->d_hpart, ctx->gridvol,
                               d_H[c]);
        }
        HIP_TRY(ctx, hipGetLastError());
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    collect_timings(ctx);
    return LCPAO_OK;
}

}  // namespace

void lcpao::launch_deff_f64(lcpao_ctx *ctx, const double *D, int sym_self) {
    hipLaunchKernelGGL(k_deff<double>, dim3((unsigned)ctx->ncanon), dim3(128), 0, ctx->stream, ctx->d_hcanon, ctx->ncanon,
                       ctx->d_pairs, D, ctx->d_deff, sym_self);
}
void lcpao::launch_deff_f32_batch(hipStream_t stream, int nstruct, const lcpao::BatchFold *d_fold,
                                  const int32_t *d_off, int total_canon) {
    if (nstruct > 0 && total_canon > 0)
        hipLaunchKernelGGL(k_deff_batch, dim3((unsigned)total_canon), dim3(128), 0, stream, d_fold, d_off, nstruct);
}
void lcpao::launch_deff_f32(lcpao_ctx *ctx, const double *D, int sym_self) {
    hipLaunchKernelGGL(k_deff<float>, dim3((unsigned)ctx->ncanon), dim3(128), 0, ctx->stream, ctx->d_hcanon, ctx->ncanon,
                       ctx->d_pairs, D, ctx->d_deff32, sym_self);
}

namespace {

// Runs `launch(variant)` once per candidate (after one untimed run) with
// events on the context stream and keeps the fastest; the caller's real
// launch follows, so every candidate writes only scratch that it overwrites.
// Candidates with available(v) false are not run (their time reads -1).
template <class F, class A>
int autotune(lcpao_ctx *ctx, int op, F launch, A available) {
    if (ctx->variant[op] >= 0) return LCPAO_OK;
    hipEvent_t e0, e1;
    HIP_TRY(ctx, hipEventCreate(&e0));
    HIP_TRY(ctx, hipEventCreate(&e1));
    int best = -1;
    float best_ms = 0.f;
    for (int v = 0; v < TUNE_NCAND[op]; v++) {
        if (!available(v)) { ctx->variant_ms[op][v] = -1.f; continue; }
        launch(v);
        HIP_TRY(ctx, hipEventRecord(e0, ctx->stream));
        launch(v);
        HIP_TRY(ctx, hipEventRecord(e1, ctx->stream));
        HIP_TRY(ctx, hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_TRY(ctx, hipEventElapsedTime(&ms, e0, e1));
        ctx->variant_ms[op][v] = ms;
        if (best < 0 || ms < best_ms) { best = v; best_ms = ms; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIP_TRY(ctx, hipGetLastError());
    ctx->variant[op] = best;
    return LCPAO_OK;
}

int rho_dev_pairs(lcpao_ctx *ctx, int nspin, const double *const *d_DM, int64_t dm_offset, double *const *d_rho,
                  bool f32, int variant);

int rho_dev(lcpao_ctx *ctx, int nspin, const double *const *d_DM, int64_t dm_offset, double *const *d_rho, int precision) {
    if (!ctx->tables) return set_err(ctx, LCPAO_ESTATE, "orbital tables are not set");
    if (nspin != 1 && nspin != 2 && nspin != 4) return set_err(ctx, LCPAO_EINVAL, "nspin must be 1, 2 or 4");
    if (dm_offset < 0) return set_err(ctx, LCPAO_EINVAL, "negative dm_offset");
    int rc = check_precision(ctx, precision);
    if (rc) return rc;
    if (precision == LCPAO_BF16X3) {
        rc = split_density(ctx, nspin, d_DM, dm_offset, d_rho);
        if (rc != LCPAO_ERANGE || ctx->block_state != 0) return rc;
        // no block plan for this geometry: the FP32 pair-list kernel serves
        // the bf16x3 density (FP32 products are within the bf16x3 bound)
        ctx->err.clear();
        precision = LCPAO_FP32;
    }
    const bool f32 = precision != LCPAO_FP64;
    for (int s = 0; s < nspin; s++)
        if (!d_DM[s] || !d_rho[s]) return set_err(ctx, LCPAO_EINVAL, "null spin pointer");
    const int op = f32 ? TUNE_RHO32 : TUNE_RHO64;
    // candidates that cannot serve this context: the FP32 block form when the
    // block plan cannot be set up (a geometry beyond it, no cell lists, no
    // memory for its tiles), the FP64 matrix-core kernel beyond 32 orbitals
    bool block_ok = false;
    if (f32 && ctx->variant[TUNE_RHO32] != 0 && ctx->variant[TUNE_RHO32] != 2 && ctx->variant[TUNE_RHO32] != 3) {
        const int r2 = split_tables(ctx, LCPAO_FP32);
        if (r2 == LCPAO_OK) {
            block_ok = true;
        } else if (ctx->variant[TUNE_RHO32] == 1 && ctx->block_state != 0) {
            return r2;  // fixed to the block form, which failed
        } else {
            if (ctx->variant[TUNE_RHO32] == 1) ctx->variant[TUNE_RHO32] = 0;  // beyond the block limits
            ctx->err.clear();
            (void)hipGetLastError();  // a failed allocation leaves no sticky state behind
        }
    }
    const bool m64_ok = ctx->maxno <= MFMA64_NO;
    if (!f32 && ctx->variant[TUNE_RHO64] == 0 && !m64_ok)
        return set_err(ctx, LCPAO_ERANGE, "the FP64 matrix-core density kernel holds at most 32 orbitals per atom");
    // the FP64 block form (candidate 6) needs the block plan
    bool block64_ok = false;
    const int v64 = ctx->variant[TUNE_RHO64];
    const bool v64_block = v64 >= RHO64_BLOCK;
    if (!f32 && (v64 < 0 || v64_block)) {
        const int r2 = plan_build(ctx);
        if (r2 == LCPAO_OK) {
            block64_ok = true;
        } else if (v64_block && ctx->block_state != 0) {
            return r2;  // fixed to the block form, which failed
        } else {
            if (v64_block) ctx->variant[TUNE_RHO64] = 1;  // beyond the block limits
            ctx->err.clear();
            (void)hipGetLastError();
        }
    }
    auto available = [&](int v) {
        if (f32) return v != 1 || block_ok;
        if (v >= RHO64_BLOCK) return block64_ok;
        return v != 0 || m64_ok;
    };
    auto run = [&](int v) -> int {
        if (f32 && v == 1) return split_density_f32_enqueue(ctx, nspin, d_DM, dm_offset, d_rho);
        if (!f32 && v >= RHO64_BLOCK)
            return split_density_f64_enqueue(ctx, nspin, d_DM, dm_offset, d_rho, v - RHO64_BLOCK);
        return rho_dev_pairs(ctx, nspin, d_DM, dm_offset, d_rho, f32, v);
    };
    if (f32 && (ctx->variant[op] != 1) && (rc = fp32_tables(ctx, false))) return rc;
    if (ctx->variant[op] < 0) {
        // every candidate overwrites rho; tuned with every spin's work
        int trc = LCPAO_OK;
        auto cand = [&](int v) {
            if (!trc) trc = run(v);
        };
        if ((rc = autotune(ctx, op, cand, available))) return rc;
        if (trc) return trc;
        if (f32 && ctx->variant[op] == 1) {
            // the pair-list candidates' FP32 table copy is no longer needed
            // (fp32_tables makes it again for the FP32 Force3)
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            if (ctx->d_tab32) { (void)hipFree(ctx->d_tab32); ctx->d_tab32 = nullptr; }
        }
    }
    if ((rc = run(ctx->variant[op]))) return rc;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    collect_timings(ctx);
    return LCPAO_OK;
}

// The pair-list density kernels (FP64 candidates 0 and 2..5 or the FP32
// ones 0, 2, 3) for every spin: fold, contraction into the per-entry
// scratch, gather into rho (enqueue only).
int rho_dev_pairs(lcpao_ctx *ctx, int nspin, const double *const *d_DM, int64_t dm_offset, double *const *d_rho,
                  bool f32, int variant) {
    // (kernel, waves, decomposition) of the variant
    int kind = 1, waves = 4, dec = 0;  // kind 0: matrix cores, 1: vector FMA
    if (f32) {
        if (variant == 2) { dec = 1; }
        else if (variant == 3) { waves = 8; dec = 2; }
    } else {
        switch (variant) {
            case 0: kind = 0; break;
            case 2: dec = 1; break;
            case 3: waves = 8; dec = 2; break;
            case 4: waves = 2; dec = 1; break;
            case 5: dec = 2; break;
            default: break;
        }
    }
    const auto &R = ctx->rdec[dec];
    const int nu = (int)R.nunits;
    for (int s = 0; s < nspin; s++) {
        const double *D = d_DM[s] + dm_offset;
        if (ctx->ncanon > 0) {
            KTimer kt(ctx, KID_DEFF);
            if (f32)
                hipLaunchKernelGGL(k_deff<float>, dim3((unsigned)ctx->ncanon), dim3(128), 0, ctx->stream, ctx->d_hcanon,
                                   ctx->ncanon, ctx->d_pairs, D, ctx->d_deff32);
            else
                hipLaunchKernelGGL(k_deff<double>, dim3((unsigned)ctx->ncanon), dim3(128), 0, ctx->stream, ctx->d_hcanon,
                                   ctx->ncanon, ctx->d_pairs, D, ctx->d_deff);
        }
        if (nu > 0) {
            KTimer kt(ctx, KID_RHO);
#define RHO_ARGS R.d_units, R.d_order, nu, R.d_segs, ctx->d_pairs
#define RHO_TAIL ctx->d_olg1, ctx->d_olg2
            if (f32) {
                if (waves == 8)
                    hipLaunchKernelGGL((k_rho<float, 8, 512>), dim3(nu), dim3(64 * 8), 0, ctx->stream, RHO_ARGS, ctx->d_tab32,
                                       RHO_TAIL, ctx->d_deff32, ctx->d_gl_off, ctx->d_rhoA);
                else if (dec == 1)
                    hipLaunchKernelGGL((k_rho<float, 4, 1024>), dim3(nu), dim3(64 * 4), 0, ctx->stream, RHO_ARGS, ctx->d_tab32,
                                       RHO_TAIL, ctx->d_deff32, ctx->d_gl_off, ctx->d_rhoA);
                else
                    hipLaunchKernelGGL((k_rho<float, 4, 2048>), dim3(nu), dim3(64 * 4), 0, ctx->stream, RHO_ARGS, ctx->d_tab32,
                                       RHO_TAIL, ctx->d_deff32, ctx->d_gl_off, ctx->d_rhoA);
            } else if (kind == 0) {
                hipLaunchKernelGGL(k_rho_m64, dim3(nu), dim3(64 * RHO_WAVES), 0, ctx->stream, RHO_ARGS, ctx->d_tab,
                                   RHO_TAIL, ctx->d_deff, ctx->d_gl_off, ctx->d_rhoA);
            } else {
#define RHO64(W, CH)                                                                                              \
    hipLaunchKernelGGL((k_rho<double, W, CH>), dim3(nu), dim3(64 * W), 0, ctx->stream, RHO_ARGS, ctx->d_tab, RHO_TAIL, \
                       ctx->d_deff, ctx->d_gl_off, ctx->d_rhoA)
                if (waves == 4 && dec == 0) RHO64(4, 2048);
                else if (waves == 4 && dec == 1) RHO64(4, 1024);
                else if (waves == 8 && dec == 2) RHO64(8, 512);
                else if (waves == 2 && dec == 1) RHO64(2, 1024);
                else RHO64(4, 512);
#undef RHO64
            }
#undef RHO_ARGS
#undef RHO_TAIL
        }
        const unsigned gb = (unsigned)((ctx->nn + 255) / 256);
        {
            KTimer kt(ctx, KID_RGATHER);
            hipLaunchKernelGGL(k_rho_gather, dim3(gb), dim3(256), 0, ctx->stream, ctx->d_ent_off, ctx->d_ent, ctx->d_rhoA,
                               ctx->nn, d_rho[s]);
        }
        HIP_TRY(ctx, hipGetLastError());
    }
    return LCPAO_OK;
}

int f3_dev(lcpao_ctx *ctx, int nch, const double *const *d_V, const double *const *d_DM, int64_t dm_offset,
           const double *weight, double *d_force, int precision) {
    if (!ctx->tables || !ctx->dtables) return set_err(ctx, LCPAO_ESTATE, "orbital and gradient tables must be set");
    if (nch != 1 && nch != 2 && nch != 4) return set_err(ctx, LCPAO_EINVAL, "nch must be 1, 2 or 4");
    if (!weight || !d_force) return set_err(ctx, LCPAO_EINVAL, "null weight or force pointer");
    int rc = check_precision(ctx, precision);
    if (rc) return rc;
    if (precision == LCPAO_BF16X3 && (rc = split_force3_ready(ctx))) {
        if (rc != LCPAO_ERANGE) return rc;
        // the block form cannot hold this geometry: the FP32 pair-list
        // kernel serves the bf16x3 Force3 (FP32 products are within its bound)
        ctx->err.clear();
        precision = LCPAO_FP32;
    }
    const bool bf = precision == LCPAO_BF16X3;
    const bool f32 = precision == LCPAO_FP32;
    if (!f32 && !bf && ctx->maxno > MFMA64_NO) {  // the matrix-core candidate holds at most 32 orbitals
        if (ctx->variant[TUNE_F364] == 0) return set_err(ctx, LCPAO_ERANGE, "the FP64 matrix-core Force3 kernel holds at most 32 orbitals per atom");
        ctx->variant[TUNE_F364] = 1;
    }
    if (f32 && (rc = fp32_tables(ctx, true))) return rc;
    const int nu = (int)ctx->funits.size();
    if (ctx->natom > 0)
        hipLaunchKernelGGL(k_fill_zero, dim3((ctx->natom * 3 + 255) / 256), dim3(256), 0, ctx->stream,
                           (int64_t)ctx->natom * 3, ctx->d_fsum);
    for (int c = 0; c < nch; c++) {
        if (!d_V[c] || !d_DM[c]) return set_err(ctx, LCPAO_EINVAL, "null channel pointer");
        const double *D = d_DM[c] + dm_offset;
        if (bf) {  // block form on the matrix cores
            if ((rc = split_force3(ctx, d_V[c], D, weight[c], ctx->d_fsum))) return rc;
            continue;
        }
        if (f32) launch_to_f32(ctx->stream, ctx->my_hsize, D, ctx->d_deff32);
        // FP64 candidates: 0 = matrix cores (k_f3_m64), 1 = vector FMA (k_f3<double>)
        auto f364 = [&](int v) {
            if (v == 0)
                hipLaunchKernelGGL(k_f3_m64, dim3(ctx->forder_n), dim3(64), 0, ctx->stream, ctx->d_funits, ctx->d_forder,
                                   ctx->forder_n,
                                   ctx->d_pairs, ctx->d_tab, ctx->d_dtab[0], ctx->d_dtab[1], ctx->d_dtab[2],
                                   ctx->d_olg1, ctx->d_olg2, ctx->d_glist, d_V[c], D, ctx->d_fpart);
            else
                hipLaunchKernelGGL(k_f3<double>, dim3(ctx->forder_n), dim3(64), 0, ctx->stream, ctx->d_funits, ctx->d_forder,
                                   ctx->forder_n,
                                   ctx->d_pairs, ctx->d_tab, ctx->d_dtab[0], ctx->d_dtab[1], ctx->d_dtab[2],
                                   ctx->d_olg1, ctx->d_olg2, ctx->d_glist, d_V[c], D, ctx->d_fpart);
        };
        if (nu > 0 && !f32 && (rc = autotune(ctx, TUNE_F364, f364, [&](int v) { return v != 0 || ctx->maxno <= MFMA64_NO; })))
            return rc;
        if (nu > 0) {
            KTimer kt(ctx, KID_F3);
            if (f32)
                hipLaunchKernelGGL(k_f3<float>, dim3(ctx->forder_n), dim3(64), 0, ctx->stream, ctx->d_funits, ctx->d_forder,
                                   ctx->forder_n,
                                   ctx->d_pairs, ctx->d_tab32, ctx->d_dtab32[0], ctx->d_dtab32[1], ctx->d_dtab32[2],
                                   ctx->d_olg1, ctx->d_olg2, ctx->d_glist, d_V[c], ctx->d_deff32, ctx->d_fpart);
            else
                f364(ctx->variant[TUNE_F364]);
        }
        if (nu > 0) {
            KTimer kt(ctx, KID_F3RED);
            hipLaunchKernelGGL(k_f3_reduce, dim3((ctx->natom + 3) / 4), dim3(256), 0, ctx->stream, ctx->d_atom_ub,
                               ctx->natom, ctx->d_fpart, weight[c], ctx->d_fsum);
        }
        HIP_TRY(ctx, hipGetLastError());
    }
    hipLaunchKernelGGL(k_f3_apply, dim3((ctx->natom + 127) / 128), dim3(128), 0, ctx->stream, ctx->natom, ctx->d_fsum,
                       ctx->gridvol, d_force);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    collect_timings(ctx);
    return LCPAO_OK;
}

// FP64 Force3 with its per-unit partials and per-atom sums in the
// file-scope scratch array instead of the context's buffers: the same
// kernels in the same order as f3_dev (so the same bits for the same kernel
// variant), no allocation, no event timing and no autotuning (a context
// whose FP64 Force3 variant is still open runs the vector-FMA kernel).
int f3_dev_noalloc(lcpao_ctx *ctx, int nch, const double *const *d_V, const double *const *d_DM, int64_t dm_offset,
                   const double *weight, double *d_force) {
    if (!ctx->tables || !ctx->dtables) return set_err(ctx, LCPAO_ESTATE, "orbital and gradient tables must be set");
    if (nch != 1 && nch != 2 && nch != 4) return set_err(ctx, LCPAO_EINVAL, "nch must be 1, 2 or 4");
    if (!weight || !d_force) return set_err(ctx, LCPAO_EINVAL, "null weight or force pointer");
    if (dm_offset < 0) return set_err(ctx, LCPAO_EINVAL, "negative dm_offset");
    for (int c = 0; c < nch; c++)
        if (!d_V[c] || !d_DM[c]) return set_err(ctx, LCPAO_EINVAL, "null channel pointer");
    double *base = nullptr;
    hipError_t src = hipErrorInvalidDevice;
    if (ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES) {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (g_scratch_done[ctx->device]) { base = g_scratch_ptr[ctx->device]; src = g_scratch_rc[ctx->device]; }
    }
    if (!base)
        return set_err(ctx, LCPAO_EDEVICE, std::string("scratch address lookup failed: ") + hipGetErrorString(src));
    const int nu = (int)ctx->funits.size();
    const int64_t fpart_n = ((3 * (int64_t)nu + 31) / 32) * 32;  // keep the per-atom sums 256-byte aligned
    const int64_t need = fpart_n + 3 * (int64_t)ctx->natom;
    if (need > NOALLOC_SCRATCH_DOUBLES)
        return set_err(ctx, LCPAO_ERANGE, "geometry needs " + std::to_string(need) + " doubles of scratch (" +
                                              std::to_string(nu) + " Force3 units, " + std::to_string(ctx->natom) +
                                              " atoms); the allocation-free entry point holds at most " +
                                              std::to_string((long long)NOALLOC_SCRATCH_DOUBLES));
    double *fpart = base, *fsum = base + fpart_n;
    const int variant = ctx->variant[TUNE_F364] >= 0 ? ctx->variant[TUNE_F364] : 1;
    if (variant == 0 && ctx->maxno > MFMA64_NO)
        return set_err(ctx, LCPAO_ERANGE, "the FP64 matrix-core Force3 kernel holds at most 32 orbitals per atom");
    if (ctx->natom > 0)
        hipLaunchKernelGGL(k_fill_zero, dim3((ctx->natom * 3 + 255) / 256), dim3(256), 0, ctx->stream,
                           (int64_t)ctx->natom * 3, fsum);
    for (int c = 0; c < nch; c++) {
        const double *D = d_DM[c] + dm_offset;
        if (nu > 0) {
            if (variant == 0)
                hipLaunchKernelGGL(k_f3_m64, dim3(ctx->forder_n), dim3(64), 0, ctx->stream, ctx->d_funits, ctx->d_forder,
                                   ctx->forder_n,
                                   ctx->d_pairs, ctx->d_tab, ctx->d_dtab[0], ctx->d_dtab[1], ctx->d_dtab[2],
                                   ctx->d_olg1, ctx->d_olg2, ctx->d_glist, d_V[c], D, fpart);
            else
                hipLaunchKernelGGL(k_f3<double>, dim3(ctx->forder_n), dim3(64), 0, ctx->stream, ctx->d_funits, ctx->d_forder,
                                   ctx->forder_n,
                                   ctx->d_pairs, ctx->d_tab, ctx->d_dtab[0], ctx->d_dtab[1], ctx->d_dtab[2],
                                   ctx->d_olg1, ctx->d_olg2, ctx->d_glist, d_V[c], D, fpart);
            hipLaunchKernelGGL(k_f3_reduce, dim3((ctx->natom + 3) / 4), dim3(256), 0, ctx->stream, ctx->d_atom_ub,
                               ctx->natom, fpart, weight[c], fsum);
        }
        HIP_TRY(ctx, hipGetLastError());
    }
    hipLaunchKernelGGL(k_f3_apply, dim3((ctx->natom + 127) / 128), dim3(128), 0, ctx->stream, ctx->natom, fsum,
                       ctx->gridvol, d_force);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return LCPAO_OK;
}

void free_all(lcpao_ctx *ctx) {
    void *ptrs[] = {ctx->d_glist, ctx->d_clist, ctx->d_gl_off, ctx->d_tab_off, ctx->d_no, ctx->d_olg1, ctx->d_olg2,
                    ctx->d_pairs, ctx->d_tab, ctx->d_dtab[0], ctx->d_dtab[1], ctx->d_dtab[2], ctx->d_hunits,
                    ctx->d_horder, ctx->d_hpair_ub, ctx->d_hcanon, ctx->d_hpart, ctx->d_deff, ctx->d_rhoA,
                    ctx->d_ent_off, ctx->d_ent, ctx->d_funits, ctx->d_forder,
                    ctx->d_atom_ub, ctx->d_fpart, ctx->d_fsum, ctx->d_stage, ctx->d_flag};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (auto &D : ctx->rdec) {
        if (D.d_units) (void)hipFree(D.d_units);
        if (D.d_order) (void)hipFree(D.d_order);
        if (D.d_segs) (void)hipFree(D.d_segs);
    }
    invalidate_split_tables(ctx);
    if (ctx->d_deff32) (void)hipFree(ctx->d_deff32);
    plan_free(ctx->plan);
    ctx->plan = nullptr;
    field_free(ctx->fields);
    ctx->fields = nullptr;
    bloch_free(ctx->bloch);
    ctx->bloch = nullptr;
    for (auto &p : ctx->pending) { (void)hipEventDestroy(p.beg); (void)hipEventDestroy(p.end); }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->stream && ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
}

}  // namespace

namespace lcpao {
int ctx_enter(lcpao_ctx *ctx) { return check_ctx(ctx); }
int ctx_error(lcpao_ctx *ctx, int code, const std::string &msg) { return set_err(ctx, code, msg); }
void cell_triplet(const lcpao_ctx *ctx, int cell, int *ijk) { cell_ijk(ctx->cpy, cell, ijk); }
std::vector<double> cell_vectors(const lcpao_ctx *ctx) {
    const int C = ctx->cpy, W = 2 * C + 1, ncell = W * W * W;
    std::vector<double> atv((size_t)ncell * 3, 0.0);
    for (int c = 1; c < ncell; c++) {
        int ijk[3];
        cell_ijk(C, c, ijk);
        const double di = ijk[0], dj = ijk[1], dk = ijk[2];
        for (int x = 0; x < 3; x++) atv[3 * (size_t)c + x] = di * ctx->lat[x] + dj * ctx->lat[3 + x] + dk * ctx->lat[6 + x];
    }
    return atv;
}
}  // namespace lcpao

extern "C" {

int lcpao_grid_create(lcpao_ctx **out, const lcpao_geom *g, int device) {
    g_create_err.clear();
    if (!out || !g) { g_create_err = "null argument"; return LCPAO_EINVAL; }
    *out = nullptr;
    if (g->natom <= 0 || !g->no || !g->gridn || !g->grid_list || g->npairs < 0 ||
        (g->npairs > 0 && (!g->pair_atom || !g->pair_natn || !g->pair_ncn || !g->numolg || !g->olg1 || !g->olg2)) ||
        g->ngrid[0] <= 0 || g->ngrid[1] <= 0 || g->ngrid[2] <= 0 || g->cpycell < 0 ||
        (g->atom_base != 0 && g->atom_base != 1)) {
        g_create_err = "invalid geometry description";
        return LCPAO_EINVAL;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) { g_create_err = "no HIP device available"; return LCPAO_EDEVICE; }
    auto *ctx = new lcpao_ctx();
    if (device < 0) { if (hipGetDevice(&device) != hipSuccess) device = 0; }
    if (device >= ndev) { g_create_err = "device index out of range"; delete ctx; return LCPAO_EDEVICE; }
    ctx->device = device;
    auto fail = [&](int code, const std::string &m) {
        g_create_err = m.empty() ? ctx->err : m;
        free_all(ctx);
        delete ctx;
        return code;
    };
    if (hipSetDevice(device) != hipSuccess) return fail(LCPAO_EDEVICE, "hipSetDevice failed");
    scratch_setup(device);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return fail(LCPAO_EDEVICE, "stream creation failed");
    ctx->natom = g->natom;
    ctx->cpy = g->cpycell;
    for (int d = 0; d < 3; d++) ctx->ngrid[d] = g->ngrid[d];
    ctx->nn = (int64_t)g->ngrid[0] * g->ngrid[1] * g->ngrid[2];
    if (ctx->nn > INT32_MAX) return fail(LCPAO_ERANGE, "grid larger than 2^31 points");
    ctx->gridvol = g->gridvol;
    ctx->no.assign(g->no, g->no + g->natom);
    ctx->gridn.assign(g->gridn, g->gridn + g->natom);
    ctx->gl_off.assign(g->natom + 1, 0);
    ctx->tab_off.assign(g->natom + 1, 0);
    for (int a = 0; a < g->natom; a++) {
        if (ctx->no[a] < 0 || ctx->no[a] > MAXNO) return fail(LCPAO_ERANGE, "orbital count per atom must be in [0, 144]");
        if (ctx->gridn[a] < 0 || (ctx->gridn[a] > 0 && !g->grid_list[a])) return fail(LCPAO_EINVAL, "bad grid list");
        ctx->maxno = std::max(ctx->maxno, ctx->no[a]);
        ctx->gl_off[a + 1] = ctx->gl_off[a] + ctx->gridn[a];
        ctx->tab_off[a + 1] = ctx->tab_off[a] + (int64_t)ctx->gridn[a] * ctx->no[a];
    }
    ctx->total_gridn = ctx->gl_off[g->natom];
    std::vector<int32_t> glist(ctx->total_gridn), clist;
    for (int a = 0; a < g->natom; a++)
        for (int t = 0; t < ctx->gridn[a]; t++) {
            const int32_t v = g->grid_list[a][t];
            if (v < 0 || v >= ctx->nn) return fail(LCPAO_EINVAL, "grid index out of range");
            glist[ctx->gl_off[a] + t] = v;
        }
    if (g->cell_list) {
        const int ncell = (2 * g->cpycell + 1) * (2 * g->cpycell + 1) * (2 * g->cpycell + 1);
        clist.resize(ctx->total_gridn);
        for (int a = 0; a < g->natom; a++)
            for (int t = 0; t < ctx->gridn[a]; t++) {
                const int32_t v = g->cell_list[a][t];
                if (v < 0 || v >= ncell) return fail(LCPAO_EINVAL, "cell index out of range");
                clist[ctx->gl_off[a] + t] = v;
            }
    }
    // pairs
    ctx->npair = g->npairs;
    ctx->pairs.resize(g->npairs);
    int64_t hoff = 0, olg = 0;
    const int ncell = (2 * g->cpycell + 1) * (2 * g->cpycell + 1) * (2 * g->cpycell + 1);
    for (int64_t p = 0; p < g->npairs; p++) {
        const int a = g->pair_atom[p] - g->atom_base, b = g->pair_natn[p] - g->atom_base;
        if (a < 0 || a >= g->natom || b < 0 || b >= g->natom) return fail(LCPAO_EINVAL, "pair atom index out of range");
        if (g->pair_ncn[p] < 0 || g->pair_ncn[p] >= ncell) return fail(LCPAO_EINVAL, "pair cell index out of range");
        if (g->numolg[p] < 0 || (g->numolg[p] > 0 && (!g->olg1[p] || !g->olg2[p]))) return fail(LCPAO_EINVAL, "bad overlap list");
        PairDev &P = ctx->pairs[p];
        P.a = a; P.b = b; P.noa = ctx->no[a]; P.nob = ctx->no[b];
        P.hoff = hoff; P.olg = olg; P.tab_a = ctx->tab_off[a]; P.tab_b = ctx->tab_off[b];
        P.gl_a = (int32_t)ctx->gl_off[a]; P.partner = -1;
        if (ctx->gl_off[a] > INT32_MAX) return fail(LCPAO_ERANGE, "grid lists longer than 2^31 entries");
        hoff += (int64_t)P.noa * P.nob;
        olg += g->numolg[p];
    }
    ctx->my_hsize = hoff;
    ctx->total_olg = olg;
    // overlap lists, each pair's list ordered by the home-atom index t1
    std::vector<int32_t> o1(olg), o2(olg);
    for (int64_t p = 0; p < g->npairs; p++) {
        const int n = g->numolg[p];
        const int32_t *L1 = g->olg1[p], *L2 = g->olg2[p];
        const PairDev &P = ctx->pairs[p];
        std::vector<int32_t> idx(n);
        std::iota(idx.begin(), idx.end(), 0);
        bool sorted = true;
        for (int k = 0; k < n; k++) {
            if (L1[k] < 0 || L1[k] >= ctx->gridn[P.a] || L2[k] < 0 || L2[k] >= ctx->gridn[P.b])
                return fail(LCPAO_EINVAL, "overlap index out of range");
            if (k > 0 && L1[k] < L1[k - 1]) sorted = false;
        }
        if (!sorted) std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return L1[x] < L1[y]; });
        for (int k = 0; k < n; k++) { o1[P.olg + k] = L1[idx[k]]; o2[P.olg + k] = L2[idx[k]]; }
    }
    ctx->pair_ncn.assign(g->pair_ncn, g->pair_ncn + g->npairs);
    ctx->numolg.assign(g->numolg, g->numolg + g->npairs);
    int rc;
    if ((rc = upload(ctx, &ctx->d_glist, glist))) return fail(rc, "");
    if (!clist.empty() && (rc = upload(ctx, &ctx->d_clist, clist))) return fail(rc, "");
    if ((rc = upload(ctx, &ctx->d_gl_off, ctx->gl_off))) return fail(rc, "");
    if ((rc = upload(ctx, &ctx->d_tab_off, ctx->tab_off))) return fail(rc, "");
    if ((rc = upload(ctx, &ctx->d_no, ctx->no))) return fail(rc, "");
    if ((rc = upload(ctx, &ctx->d_olg1, o1))) return fail(rc, "");
    if ((rc = upload(ctx, &ctx->d_olg2, o2))) return fail(rc, "");
    if ((rc = build_work(ctx, g, o1, o2, glist))) return fail(rc, "");
    ctx->h_glist = std::move(glist);
    ctx->h_clist = std::move(clist);
    if ((rc = upload(ctx, &ctx->d_pairs, ctx->pairs))) return fail(rc, "");
    if ((rc = dev_alloc(ctx, (void **)&ctx->d_flag, sizeof(int32_t)))) return fail(rc, "");
    if (g->latvecs && g->origin && g->gxyz && g->atom2spe && g->cell_list) {
        ctx->have_lattice = true;
        std::memcpy(ctx->lat, g->latvecs, sizeof ctx->lat);
        std::memcpy(ctx->origin, g->origin, sizeof ctx->origin);
        ctx->gxyz.assign(g->gxyz, g->gxyz + 3 * (size_t)g->natom);
        ctx->atom2spe.assign(g->atom2spe, g->atom2spe + g->natom);
    }
    *out = ctx;
    return LCPAO_OK;
}

void lcpao_grid_destroy(lcpao_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    free_all(ctx);
    delete ctx;
}

int lcpao_grid_set_stream(lcpao_ctx *ctx, void *hip_stream) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->own_stream) {
        if (hip_stream) { (void)hipStreamDestroy(ctx->stream); ctx->stream = (hipStream_t)hip_stream; ctx->own_stream = false; }
    } else {
        if (hip_stream) ctx->stream = (hipStream_t)hip_stream;
        else { HIP_TRY(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)); ctx->own_stream = true; }
    }
    return LCPAO_OK;
}

int lcpao_grid_profile(lcpao_ctx *ctx, int32_t enable) {
    if (!ctx) return LCPAO_EINVAL;
    ctx->profiling = enable != 0;
    for (int k = 0; k < KID_COUNT; k++) { ctx->prof_launches[k] = 0; ctx->prof_ms[k] = 0.0; }
    return LCPAO_OK;
}

int lcpao_grid_tune(lcpao_ctx *ctx, int32_t op, int32_t variant) {
    if (!ctx) return LCPAO_EINVAL;
    if (op < 0 || op >= TUNE_NOPS) return set_err(ctx, LCPAO_EINVAL, "unknown tunable operation");
    if (variant < -1 || variant >= TUNE_NCAND[op]) return set_err(ctx, LCPAO_EINVAL, "variant out of range");
    ctx->variant[op] = variant;
    return LCPAO_OK;
}

int lcpao_grid_tuned(const lcpao_ctx *ctx, int32_t op, int32_t *variant, float *candidate_ms, int32_t ncand) {
    if (!ctx || !variant || op < 0 || op >= TUNE_NOPS) return LCPAO_EINVAL;
    *variant = ctx->variant[op];
    for (int v = 0; v < ncand && v < TUNE_NCAND[op]; v++)
        if (candidate_ms) candidate_ms[v] = ctx->variant_ms[op][v];
    return LCPAO_OK;
}

}  // extern "C"

namespace {

// Key of a context's shape for the tuner cache: the distinct orbital counts
// of its atoms, the grid spacing (cube root of GridVol) and the mean number
// of pairs per atom, which set the kernels' tile shapes and work per pair.
std::string tune_key(const lcpao_ctx *ctx) {
    std::vector<int> nos(ctx->no.begin(), ctx->no.end());
    std::sort(nos.begin(), nos.end());
    nos.erase(std::unique(nos.begin(), nos.end()), nos.end());
    std::string k = "no=";
    for (size_t i = 0; i < nos.size(); i++) k += (i ? "," : "") + std::to_string(nos[i]);
    char buf[96];
    std::snprintf(buf, sizeof buf, ";h=%.4f;ppa=%lld", std::cbrt(ctx->gridvol),
                  (long long)(ctx->natom > 0 ? (ctx->npair + ctx->natom / 2) / ctx->natom : 0));
    return k + buf;
}

struct TuneEntry {
    std::string key;
    int op, variant;
    std::vector<float> ms;   // measured ms of each candidate (-1: not run)
};

// The cache is a JSON object with one entry per line:
//   {"key": "no=19;h=0.2225;ppa=87", "op": 0, "variant": 1, "ms": [19.9, 11.9, ...]}
// Lines that do not hold an entry are skipped.  Returns false if the file
// cannot be opened.
bool read_tune_file(const char *path, std::vector<TuneEntry> &out) {
    FILE *f = std::fopen(path, "r");
    if (!f) return false;
    char line[1024];
    while (std::fgets(line, sizeof line, f)) {
        const char *p = std::strstr(line, "{\"key\": \"");
        if (!p) continue;
        char key[512];
        int op = -1, v = -1, used = 0;
        if (std::sscanf(p, "{\"key\": \"%511[^\"]\", \"op\": %d, \"variant\": %d%n", key, &op, &v, &used) < 3) continue;
        TuneEntry e{key, op, v, {}};
        const char *m = std::strstr(p + used, "\"ms\": [");
        if (m) {
            m += 7;
            for (int k = 0; k < TUNE_MAXCAND; k++) {
                char *end = nullptr;
                const float x = std::strtof(m, &end);
                if (end == m) break;
                e.ms.push_back(x);
                m = end;
                while (*m == ' ' || *m == ',') m++;
                if (*m == ']') break;
            }
        }
        out.push_back(e);
    }
    std::fclose(f);
    return true;
}

}  // namespace

extern "C" {

int lcpao_grid_tune_load(lcpao_ctx *ctx, const char *path, int32_t *loaded) {
    if (!ctx || !path) return LCPAO_EINVAL;
    if (loaded) *loaded = 0;
    std::vector<TuneEntry> entries;
    if (!read_tune_file(path, entries)) return LCPAO_OK;  // no cache yet
    const std::string key = tune_key(ctx);
    int n = 0;
    for (const TuneEntry &e : entries) {
        if (e.key != key || e.op < 0 || e.op >= TUNE_NOPS || e.variant < 0 || e.variant >= TUNE_NCAND[e.op]) continue;
        ctx->variant[e.op] = e.variant;
        for (int v = 0; v < (int)e.ms.size() && v < TUNE_NCAND[e.op]; v++) ctx->variant_ms[e.op][v] = e.ms[v];
        n++;
    }
    if (loaded) *loaded = n;
    return LCPAO_OK;
}

int lcpao_grid_tune_save(lcpao_ctx *ctx, const char *path) {
    if (!ctx || !path) return LCPAO_EINVAL;
    std::vector<TuneEntry> entries;
    read_tune_file(path, entries);
    const std::string key = tune_key(ctx);
    std::vector<TuneEntry> keep;
    for (const TuneEntry &e : entries)
        if (!(e.key == key && e.op >= 0 && e.op < TUNE_NOPS && ctx->variant[e.op] >= 0)) keep.push_back(e);
    for (int op = 0; op < TUNE_NOPS; op++)
        if (ctx->variant[op] >= 0)
            keep.push_back(TuneEntry{key, op, ctx->variant[op],
                                     std::vector<float>(ctx->variant_ms[op], ctx->variant_ms[op] + TUNE_NCAND[op])});
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "w");
    if (!f) return set_err(ctx, LCPAO_EINVAL, "cannot write the tuner cache " + tmp);
    std::fprintf(f, "{\"lcpao_grid_tune\": [\n");
    for (size_t i = 0; i < keep.size(); i++) {
        std::fprintf(f, "{\"key\": \"%s\", \"op\": %d, \"variant\": %d, \"ms\": [", keep[i].key.c_str(), keep[i].op,
                     keep[i].variant);
        for (size_t k = 0; k < keep[i].ms.size(); k++) std::fprintf(f, "%s%.4f", k ? ", " : "", keep[i].ms[k]);
        std::fprintf(f, "]}%s\n", i + 1 < keep.size() ? "," : "");
    }
    std::fprintf(f, "]}\n");
    const bool ok = std::fclose(f) == 0;
    if (!ok || std::rename(tmp.c_str(), path) != 0) return set_err(ctx, LCPAO_EINVAL, std::string("cannot write ") + path);
    return LCPAO_OK;
}

int lcpao_grid_block_forms(const lcpao_ctx *ctx, int32_t *state) {
    if (!ctx || !state) return LCPAO_EINVAL;
    *state = ctx->block_state;
    return LCPAO_OK;
}

int lcpao_grid_kernel_stats(const lcpao_ctx *ctx, int32_t kernel_id, int64_t *launches, double *total_ms) {
    if (!ctx || kernel_id < 0 || kernel_id >= KID_COUNT) return LCPAO_EINVAL;
    if (launches) *launches = ctx->prof_launches[kernel_id];
    if (total_ms) *total_ms = ctx->prof_ms[kernel_id];
    return LCPAO_OK;
}

int lcpao_grid_sizes(const lcpao_ctx *ctx, int64_t *my_hsize, int64_t *nn, int64_t *total_olg, int64_t *total_gridn) {
    if (!ctx) return LCPAO_EINVAL;
    if (my_hsize) *my_hsize = ctx->my_hsize;
    if (nn) *nn = ctx->nn;
    if (total_olg) *total_olg = ctx->total_olg;
    if (total_gridn) *total_gridn = ctx->total_gridn;
    return LCPAO_OK;
}

static int upload_tables(lcpao_ctx *ctx, double **dst, const double *const *src) {
    if (!src) return set_err(ctx, LCPAO_EINVAL, "null table array");
    const int64_t total = ctx->tab_off[ctx->natom];
    if (!*dst) {
        int rc = dev_alloc(ctx, (void **)dst, (size_t)std::max<int64_t>(1, total) * sizeof(double));
        if (rc) return rc;
    }
    for (int a = 0; a < ctx->natom; a++) {
        const int64_t n = ctx->tab_off[a + 1] - ctx->tab_off[a];
        if (n == 0) continue;
        if (!src[a]) return set_err(ctx, LCPAO_EINVAL, "null table of an atom");
        HIP_TRY(ctx, hipMemcpyAsync(*dst + ctx->tab_off[a], src[a], (size_t)n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return LCPAO_OK;
}

int lcpao_grid_set_tables(lcpao_ctx *ctx, const double *const *phi) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    invalidate_split_tables(ctx);
    if ((rc = upload_tables(ctx, &ctx->d_tab, phi))) return rc;
    ctx->tables = true;
    return LCPAO_OK;
}

int lcpao_grid_set_dtables(lcpao_ctx *ctx, const double *const *dx, const double *const *dy, const double *const *dz) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    invalidate_split_tables(ctx);
    if ((rc = upload_tables(ctx, &ctx->d_dtab[0], dx))) return rc;
    if ((rc = upload_tables(ctx, &ctx->d_dtab[1], dy))) return rc;
    if ((rc = upload_tables(ctx, &ctx->d_dtab[2], dz))) return rc;
    ctx->dtables = true;
    return LCPAO_OK;
}

int lcpao_grid_set_positions(lcpao_ctx *ctx, const double *gxyz) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!gxyz) return set_err(ctx, LCPAO_EINVAL, "null positions");
    if (!ctx->have_lattice) return set_err(ctx, LCPAO_ESTATE, "geometry lacks lattice and positions");
    ctx->gxyz.assign(gxyz, gxyz + 3 * (size_t)ctx->natom);
    return LCPAO_OK;
}

int lcpao_grid_build_tables(lcpao_ctx *ctx, int32_t nspecies, const lcpao_pao *paos, int32_t with_derivatives) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->have_lattice || !ctx->d_clist) return set_err(ctx, LCPAO_ESTATE, "geometry lacks lattice, positions, species or cell lists");
    if (nspecies <= 0 || !paos) return set_err(ctx, LCPAO_EINVAL, "no species");
    for (int a = 0; a < ctx->natom; a++)
        if (ctx->atom2spe[a] < 0 || ctx->atom2spe[a] >= nspecies) return set_err(ctx, LCPAO_EINVAL, "atom2spe out of range");
    // radial data to the device
    std::vector<SpeciesDev> sd(nspecies);
    std::vector<double *> owned;
    auto cleanup = [&]() { for (double *p : owned) (void)hipFree(p); };
    for (int s = 0; s < nspecies; s++) {
        const lcpao_pao &P = paos[s];
        if (P.maxl < 0 || P.maxl > 3 || P.nmesh < 6 || !P.rv || !P.rwf) { cleanup(); return set_err(ctx, LCPAO_EINVAL, "bad PAO description"); }
        SpeciesDev &S = sd[s];
        std::memset(&S, 0, sizeof S);
        S.maxl = P.maxl; S.nmesh = P.nmesh;
        int no = 0;
        for (int l = 0; l <= P.maxl; l++) {
            if (P.num_basis[l] < 0 || P.num_basis[l] > 16) { cleanup(); return set_err(ctx, LCPAO_ERANGE, "multiplicity above 16"); }
            S.num_basis[l] = P.num_basis[l];
            no += P.num_basis[l] * (2 * l + 1);
        }
        S.no = no;
        double *d = nullptr;
        if (hipMalloc(&d, sizeof(double) * P.nmesh) != hipSuccess) { cleanup(); return set_err(ctx, LCPAO_EDEVICE, "hipMalloc failed"); }
        owned.push_back(d);
        (void)hipMemcpy(d, P.rv, sizeof(double) * P.nmesh, hipMemcpyHostToDevice);
        S.rv = d;
        for (int l = 0; l <= P.maxl; l++)
            for (int p = 0; p < P.num_basis[l]; p++) {
                const double *src = P.rwf[l * 16 + p];
                if (!src) { cleanup(); return set_err(ctx, LCPAO_EINVAL, "null radial function"); }
                double *w = nullptr;
                if (hipMalloc(&w, sizeof(double) * P.nmesh) != hipSuccess) { cleanup(); return set_err(ctx, LCPAO_EDEVICE, "hipMalloc failed"); }
                owned.push_back(w);
                (void)hipMemcpy(w, src, sizeof(double) * P.nmesh, hipMemcpyHostToDevice);
                S.rwf[l][p] = w;
            }
    }
    for (int a = 0; a < ctx->natom; a++)
        if (sd[ctx->atom2spe[a]].no != ctx->no[a]) { cleanup(); return set_err(ctx, LCPAO_EINVAL, "species orbital count differs from Total_NumOrbs"); }
    SpeciesDev *d_sd = nullptr;
    double *d_gxyz = nullptr, *d_atv = nullptr;
    int32_t *d_a2s = nullptr;
    const std::vector<double> atv = cell_vectors(ctx);
    bool ok = hipMalloc(&d_sd, sizeof(SpeciesDev) * nspecies) == hipSuccess &&
              hipMalloc(&d_gxyz, sizeof(double) * 3 * ctx->natom) == hipSuccess &&
              hipMalloc(&d_atv, sizeof(double) * atv.size()) == hipSuccess &&
              hipMalloc(&d_a2s, sizeof(int32_t) * ctx->natom) == hipSuccess;
    if (ok) {
        (void)hipMemcpy(d_sd, sd.data(), sizeof(SpeciesDev) * nspecies, hipMemcpyHostToDevice);
        (void)hipMemcpy(d_gxyz, ctx->gxyz.data(), sizeof(double) * 3 * ctx->natom, hipMemcpyHostToDevice);
        (void)hipMemcpy(d_atv, atv.data(), sizeof(double) * atv.size(), hipMemcpyHostToDevice);
        (void)hipMemcpy(d_a2s, ctx->atom2spe.data(), sizeof(int32_t) * ctx->natom, hipMemcpyHostToDevice);
    }
    const int64_t total = ctx->tab_off[ctx->natom];
    if (ok && !ctx->d_tab) ok = hipMalloc(&ctx->d_tab, sizeof(double) * std::max<int64_t>(1, total)) == hipSuccess;
    if (ok && with_derivatives)
        for (int x = 0; x < 3 && ok; x++)
            if (!ctx->d_dtab[x]) ok = hipMalloc(&ctx->d_dtab[x], sizeof(double) * std::max<int64_t>(1, total)) == hipSuccess;
    int result = LCPAO_OK;
    if (!ok) result = set_err(ctx, LCPAO_EDEVICE, "device allocation for table generation failed");
    if (ok) {
        TabGeom tg;
        for (int j = 0; j < 3; j++) {
            tg.glat[j] = ctx->lat[j] / ctx->ngrid[0];
            tg.glat[3 + j] = ctx->lat[3 + j] / ctx->ngrid[1];
            tg.glat[6 + j] = ctx->lat[6 + j] / ctx->ngrid[2];
            tg.origin[j] = ctx->origin[j];
        }
        tg.n2 = ctx->ngrid[1]; tg.n3 = ctx->ngrid[2]; tg.natom = ctx->natom;
        invalidate_split_tables(ctx);
        (void)hipMemsetAsync(ctx->d_flag, 0, sizeof(int32_t), ctx->stream);
        const int64_t nent = ctx->total_gridn;
        const unsigned nb = (unsigned)((nent + 127) / 128);
        if (nent > 0) {
            hipLaunchKernelGGL(k_tables, dim3(nb), dim3(128), 0, ctx->stream, nent, ctx->d_gl_off, ctx->d_glist,
                               ctx->d_clist, ctx->d_tab_off, d_gxyz, d_a2s, d_atv, tg, d_sd, 0, ctx->d_tab, nullptr,
                               nullptr, nullptr, ctx->d_flag);
            if (with_derivatives)
                hipLaunchKernelGGL(k_tables, dim3(nb), dim3(128), 0, ctx->stream, nent, ctx->d_gl_off, ctx->d_glist,
                                   ctx->d_clist, ctx->d_tab_off, d_gxyz, d_a2s, d_atv, tg, d_sd, 1, nullptr,
                                   ctx->d_dtab[0], ctx->d_dtab[1], ctx->d_dtab[2], ctx->d_flag);
        }
        int32_t flag = 0;
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(&flag, ctx->d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) result = set_err(ctx, LCPAO_EDEVICE, std::string("table generation: ") + hipGetErrorString(e));
        else if (flag) result = set_err(ctx, LCPAO_EMESH, "radial mesh index out of range (point too close to the mesh ends)");
        else {
            ctx->tables = true;
            if (with_derivatives) ctx->dtables = true;
        }
    }
    (void)hipFree(d_sd); (void)hipFree(d_gxyz); (void)hipFree(d_atv); (void)hipFree(d_a2s);
    cleanup();
    return result;
}

int lcpao_grid_get_tables(lcpao_ctx *ctx, int32_t which, double *const *out) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (which < 0 || which > 3 || !out) return set_err(ctx, LCPAO_EINVAL, "bad table selector");
    const double *src = which == 0 ? ctx->d_tab : ctx->d_dtab[which - 1];
    if (!src || (which == 0 && !ctx->tables) || (which > 0 && !ctx->dtables)) return set_err(ctx, LCPAO_ESTATE, "table not set");
    for (int a = 0; a < ctx->natom; a++) {
        const int64_t n = ctx->tab_off[a + 1] - ctx->tab_off[a];
        if (n && out[a]) HIP_TRY(ctx, hipMemcpy(out[a], src + ctx->tab_off[a], (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    }
    return LCPAO_OK;
}

int lcpao_grid_hamiltonian_dev(lcpao_ctx *ctx, int32_t nch, const double *const *d_V, double *const *d_H, int32_t precision) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!d_V || !d_H) return set_err(ctx, LCPAO_EINVAL, "null argument");
    return ham_dev(ctx, nch, d_V, d_H, precision);
}

int lcpao_grid_density_dev(lcpao_ctx *ctx, int32_t nspin, const double *const *d_DM, int64_t dm_offset,
                           double *const *d_rho, int32_t precision) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!d_DM || !d_rho) return set_err(ctx, LCPAO_EINVAL, "null argument");
    return rho_dev(ctx, nspin, d_DM, dm_offset, d_rho, precision);
}

int lcpao_grid_force3_dev(lcpao_ctx *ctx, int32_t nch, const double *const *d_V, const double *const *d_DM,
                          int64_t dm_offset, const double *weight, double *d_force, int32_t precision) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!d_V || !d_DM) return set_err(ctx, LCPAO_EINVAL, "null argument");
    return f3_dev(ctx, nch, d_V, d_DM, dm_offset, weight, d_force, precision);
}

int lcpao_grid_force3_dev_noalloc(lcpao_ctx *ctx, int32_t nch, const double *const *d_V, const double *const *d_DM,
                                  int64_t dm_offset, const double *weight, double *d_force) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!d_V || !d_DM) return set_err(ctx, LCPAO_EINVAL, "null argument");
    return f3_dev_noalloc(ctx, nch, d_V, d_DM, dm_offset, weight, d_force);
}

int lcpao_grid_hamiltonian(lcpao_ctx *ctx, int32_t nch, const double *const *V, double *const *H, int32_t precision) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!V || !H || (nch != 1 && nch != 2 && nch != 4)) return set_err(ctx, LCPAO_EINVAL, "bad channel arguments");
    const int64_t per = ctx->nn + ctx->my_hsize;
    if ((rc = stage(ctx, per * nch))) return rc;
    const double *dV[4];
    double *dH[4];
    for (int c = 0; c < nch; c++) {
        if (!V[c] || !H[c]) return set_err(ctx, LCPAO_EINVAL, "null channel pointer");
        double *v = ctx->d_stage + per * c;
        HIP_TRY(ctx, hipMemcpyAsync(v, V[c], (size_t)ctx->nn * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        dV[c] = v;
        dH[c] = v + ctx->nn;
    }
    if ((rc = ham_dev(ctx, nch, dV, dH, precision))) return rc;
    for (int c = 0; c < nch; c++)
        HIP_TRY(ctx, hipMemcpyAsync(H[c], dH[c], (size_t)ctx->my_hsize * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return LCPAO_OK;
}

int lcpao_grid_density(lcpao_ctx *ctx, int32_t nspin, const double *const *DM, int64_t dm_offset, double *const *rho,
                       int32_t precision) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!DM || !rho || (nspin != 1 && nspin != 2 && nspin != 4)) return set_err(ctx, LCPAO_EINVAL, "bad spin arguments");
    if (dm_offset < 0) return set_err(ctx, LCPAO_EINVAL, "negative dm_offset");
    const int64_t per = ctx->nn + ctx->my_hsize;
    if ((rc = stage(ctx, per * nspin))) return rc;
    const double *dD[4];
    double *dR[4];
    for (int s = 0; s < nspin; s++) {
        if (!DM[s] || !rho[s]) return set_err(ctx, LCPAO_EINVAL, "null spin pointer");
        double *d = ctx->d_stage + per * s;
        HIP_TRY(ctx, hipMemcpyAsync(d, DM[s] + dm_offset, (size_t)ctx->my_hsize * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        dD[s] = d;
        dR[s] = d + ctx->my_hsize;
    }
    if ((rc = rho_dev(ctx, nspin, dD, 0, dR, precision))) return rc;
    for (int s = 0; s < nspin; s++)
        HIP_TRY(ctx, hipMemcpyAsync(rho[s], dR[s], (size_t)ctx->nn * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return LCPAO_OK;
}

int lcpao_grid_force3(lcpao_ctx *ctx, int32_t nch, const double *const *V, const double *const *DM, int64_t dm_offset,
                      const double *weight, double *force, int32_t precision) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!V || !DM || !weight || !force || (nch != 1 && nch != 2 && nch != 4)) return set_err(ctx, LCPAO_EINVAL, "bad arguments");
    if (dm_offset < 0) return set_err(ctx, LCPAO_EINVAL, "negative dm_offset");
    const int64_t per = ctx->nn + ctx->my_hsize;
    if ((rc = stage(ctx, per * nch + 3 * (int64_t)ctx->natom))) return rc;
    const double *dV[4], *dD[4];
    for (int c = 0; c < nch; c++) {
        if (!V[c] || !DM[c]) return set_err(ctx, LCPAO_EINVAL, "null channel pointer");
        double *v = ctx->d_stage + per * c;
        HIP_TRY(ctx, hipMemcpyAsync(v, V[c], (size_t)ctx->nn * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(v + ctx->nn, DM[c] + dm_offset, (size_t)ctx->my_hsize * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        dV[c] = v;
        dD[c] = v + ctx->nn;
    }
    double *dF = ctx->d_stage + per * nch;
    HIP_TRY(ctx, hipMemcpyAsync(dF, force, 3 * (size_t)ctx->natom * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    if ((rc = f3_dev(ctx, nch, dV, dD, 0, weight, dF, precision))) return rc;
    HIP_TRY(ctx, hipMemcpyAsync(force, dF, 3 * (size_t)ctx->natom * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return LCPAO_OK;
}

const char *lcpao_grid_last_error(const lcpao_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

}  // extern "C"
