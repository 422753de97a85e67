// bloch.hip -- the steps on either side of the k-point diagonalisation:
// the Bloch sums H(k), S(k) of the flat pair blocks (HS_matrix!, and the
// 2x2 spin-block HS_matrix_NC!) and the flat pair density matrix from the
// occupied states of the k-points (Calc_DM_Crystal_Collinear_nopol!, and
// the four spin channels of Calc_DM_Crystal_NonCollinear! with the
// imaginary parts of Calc_iDM_Crystal_NonCollinear!).  Both work per atom
// pair (A, B): the images (A, B, R) of a pair share one dense block, so a
// workgroup owns one (A, B) and sums its images in the pair-list order,
// which is the order of the reference's atom / neighbour loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "ctx.hpp"

namespace lcpao {

// One workgroup's share of an (A, B) atom pair: the orbitals i0 .. i0+ni of
// A and j0 .. j0+nj of B (the whole NO_A x NO_B block up to 32 orbitals per
// atom; larger bases are cut into tiles of at most 32 x 32).
constexpr int BL_TILE = 32;
struct GroupDev {
    int32_t noa, nob;    // orbitals of A and B
    int32_t row_a, row_b;  // first row / column of A and B in the dense matrix (MP, 0-based)
    int32_t p0, np;      // the group's pairs: gpairs[p0 .. p0 + np)
    int32_t i0, j0, ni, nj;  // the tile of the NO_A x NO_B block
};
struct GPair { int64_t hoff; int32_t ijk[3], pad; };

struct BlochPlan {
    int64_t fsize = 0;
    int32_t ngroups = 0, max_nel = 0;   // ngroups: workgroup tiles
    int64_t natom_pairs = 0;            // distinct (A, B) atom pairs
    GroupDev *d_groups = nullptr;
    GPair *d_gpairs = nullptr;
    double *stage = nullptr;  // host-pointer entry points
    int64_t stage_n = 0;
};

void bloch_free(BlochPlan *b) {
    if (!b) return;
    if (b->d_groups) (void)hipFree(b->d_groups);
    if (b->d_gpairs) (void)hipFree(b->d_gpairs);
    if (b->stage) (void)hipFree(b->stage);
    delete b;
}

constexpr int BL_THREADS = 256;
constexpr int BL_MAXMAT = 4;

struct BlochArgs {
    double k[3];
    int64_t ld;                       // leading dimension of the dense matrices
    int32_t nmat, pad;
    const double *flat[BL_MAXMAT];    // real parts of the flat pair arrays
    const double *iflat[BL_MAXMAT][2];  // imaginary parts (summed, either may be null)
    int64_t off_r[BL_MAXMAT], off_c[BL_MAXMAT];  // row / column offset of the block in dense
    double2 *dense[BL_MAXMAT];
};

// HS_matrix! with k (HS_matrix.jl:88-110, 143-166): dense[m][A_i, B_j] =
// sum over the images (A, B, R) of flat[m][pair block][i, j] * cispi(2 k.R),
// R = atv_ijk[cell]; dense is column-major (ld x ld) like Julia's matrix.
// With imaginary parts (HS_matrix_NC!, HS_matrix.jl:211-257) the term is
// (flat + im*(iflat0 + iflat1)) * cispi(2 k.R), a complex product in
// Julia's order (re*c - im*s, re*s + im*c), at rows off_r + A_i and
// columns off_c + B_j.
__global__ __launch_bounds__(BL_THREADS) void k_bloch(const GroupDev *__restrict__ groups,
                                                       const GPair *__restrict__ gp, BlochArgs A) {
    const GroupDev G = groups[blockIdx.x];
    const int nel = G.ni * G.nj;
    for (int e = threadIdx.x; e < nel; e += blockDim.x) {
        const int i = G.i0 + e / G.nj, j = G.j0 + e % G.nj;
        double2 acc[BL_MAXMAT];
#pragma unroll
        for (int m = 0; m < BL_MAXMAT; m++) acc[m] = make_double2(0.0, 0.0);
        for (int q = 0; q < G.np; q++) {
            const GPair P = gp[G.p0 + q];
            const double kRn = A.k[0] * P.ijk[0] + A.k[1] * P.ijk[1] + A.k[2] * P.ijk[2];
            double sn, cs;
            sincospi(2 * kRn, &sn, &cs);
            const int64_t x = P.hoff + (int64_t)i * G.nob + j;
#pragma unroll
            for (int m = 0; m < BL_MAXMAT; m++)
                if (m < A.nmat) {
                    const double v = A.flat[m][x];
                    if (A.iflat[m][0] || A.iflat[m][1]) {
                        double w = A.iflat[m][0] ? A.iflat[m][0][x] : 0.0;
                        if (A.iflat[m][1]) w = w + A.iflat[m][1][x];
                        acc[m].x += v * cs - w * sn;
                        acc[m].y += v * sn + w * cs;
                    } else {
                        acc[m].x += v * cs;
                        acc[m].y += v * sn;
                    }
                }
        }
#pragma unroll
        for (int m = 0; m < BL_MAXMAT; m++)
            if (m < A.nmat)
                A.dense[m][(A.off_r[m] + G.row_a + i) + (A.off_c[m] + G.row_b + j) * A.ld] = acc[m];
    }
}

__global__ void k_zero_c(int64_t n, double2 *__restrict__ x) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) x[g] = make_double2(0.0, 0.0);
}

// Calc_DM_Crystal_Collinear_nopol! (Crystal_DFT.jl:219-270) without the dense
// DM: per k-point, P[j][i] = sum_mu C[rb + B_j, mu] conj(C[ra + A_i, mu]) (the
// element DM_dense[rb+Bnum+jst, ra+Anum+ist] of C C^H) for the (A, B) of the
// group only, then DM[pair][i, j] += Re(P * cispi(2 k.R) / Nk) and
// iDM[pair][i, j] += Im(P * cispi(2 k.R) / Nk) for each image, k-points in
// order.  C[k]: ld x nstates, column-major, columns already scaled by
// sqrt(occupation * weight) (Calc_fnkCnk!).  Collinear: ld = fsize, ra = rb
// = 0; the non-collinear spin blocks (Crystal_DFT.jl:342-410) take ld =
// 2 fsize and (ra, rb) = (0, 0) for up-up, (fsize, fsize) for down-down and
// (0, fsize) for the up-down block.
constexpr int DM_CHUNK = 32;  // states per LDS tile

struct DmArgs {
    int32_t nk, pad;
    int64_t ld, nstates;
    int64_t ra, rb;           // row offsets of the A and B sides in C
    double scale;             // 1 / AllNkpt
    const double *kpts;       // nk x 3
    const double2 *const *C;  // nk pointers
    double *DM;               // real part (may be null)
    double *iDM;              // imaginary part (may be null)
};

__global__ __launch_bounds__(BL_THREADS) void k_dm_states(const GroupDev *__restrict__ groups,
                                                           const GPair *__restrict__ gp, DmArgs A) {
    __shared__ double2 ta[BL_TILE][DM_CHUNK + 1];
    __shared__ double2 tb[BL_TILE][DM_CHUNK + 1];
    const GroupDev G = groups[blockIdx.x];
    const int nel = G.ni * G.nj;
    constexpr int PER = (BL_TILE * BL_TILE + BL_THREADS - 1) / BL_THREADS;
    for (int kk = 0; kk < A.nk; kk++) {
        const double2 *__restrict__ C = A.C[kk];
        double2 acc[PER];
#pragma unroll
        for (int t = 0; t < PER; t++) acc[t] = make_double2(0.0, 0.0);
        for (int64_t m0 = 0; m0 < A.nstates; m0 += DM_CHUNK) {
            const int mc = (int)min<int64_t>(DM_CHUNK, A.nstates - m0);
            __syncthreads();
            for (int x = threadIdx.x; x < (G.ni + G.nj) * DM_CHUNK; x += blockDim.x) {
                const int r = x % (G.ni + G.nj), mu = x / (G.ni + G.nj);
                double2 v = make_double2(0.0, 0.0);
                if (mu < mc) {
                    const int64_t row = r < G.ni ? A.ra + G.row_a + G.i0 + r : A.rb + G.row_b + G.j0 + (r - G.ni);
                    v = C[row + (m0 + mu) * A.ld];
                }
                if (r < G.ni) ta[r][mu] = v;
                else tb[r - G.ni][mu] = v;
            }
            __syncthreads();
#pragma unroll
            for (int t = 0; t < PER; t++) {
                const int e = threadIdx.x + t * BL_THREADS;
                if (e < nel) {
                    const int i = e / G.nj, j = e - (e / G.nj) * G.nj;
                    double re = acc[t].x, im = acc[t].y;
                    for (int mu = 0; mu < DM_CHUNK; mu++) {
                        const double2 ca = ta[i][mu], cb = tb[j][mu];
                        re += cb.x * ca.x + cb.y * ca.y;
                        im += cb.y * ca.x - cb.x * ca.y;
                    }
                    acc[t] = make_double2(re, im);
                }
            }
        }
        const double ka = A.kpts[3 * kk], kb = A.kpts[3 * kk + 1], kc = A.kpts[3 * kk + 2];
        for (int q = 0; q < G.np; q++) {
            const GPair P = gp[G.p0 + q];
            const double kRn = ka * P.ijk[0] + kb * P.ijk[1] + kc * P.ijk[2];
            double sn, cs;
            sincospi(2 * kRn, &sn, &cs);
            const double er = cs * A.scale, ei = sn * A.scale;
#pragma unroll
            for (int t = 0; t < PER; t++) {
                const int e = threadIdx.x + t * BL_THREADS;
                if (e < nel) {
                    const int64_t x = P.hoff + (int64_t)(G.i0 + e / G.nj) * G.nob + (G.j0 + e % G.nj);
                    if (A.DM) {
                        const double v = acc[t].x * er - acc[t].y * ei;
                        A.DM[x] = (kk == 0 ? 0.0 : A.DM[x]) + v;
                    }
                    if (A.iDM) {
                        const double v = acc[t].x * ei + acc[t].y * er;
                        A.iDM[x] = (kk == 0 ? 0.0 : A.iDM[x]) + v;
                    }
                }
            }
        }
    }
}

namespace {

int bloch_plan(lcpao_ctx *ctx, BlochPlan **out) {
    if (ctx->bloch) { *out = ctx->bloch; return LCPAO_OK; }
    std::vector<int64_t> mp(ctx->natom + 1, 0);
    for (int a = 0; a < ctx->natom; a++) mp[a + 1] = mp[a] + ctx->no[a];
    if (mp[ctx->natom] > INT32_MAX) return ctx_error(ctx, LCPAO_ERANGE, "too many orbitals");
    std::map<std::pair<int, int>, int> index;
    std::vector<std::vector<int>> members;
    const int64_t np = (int64_t)ctx->pairs.size();
    for (int64_t p = 0; p < np; p++) {
        const auto key = std::make_pair(ctx->pairs[p].a, ctx->pairs[p].b);
        auto it = index.find(key);
        if (it == index.end()) {
            it = index.emplace(key, (int)members.size()).first;
            members.emplace_back();
        }
        members[it->second].push_back((int)p);
    }
    std::vector<GroupDev> groups;
    std::vector<GPair> gpairs;
    int max_nel = 0;
    for (const auto &mem : members) {
        const PairDev &P0 = ctx->pairs[mem[0]];
        GroupDev g;
        g.noa = P0.noa; g.nob = P0.nob;
        g.row_a = (int32_t)mp[P0.a]; g.row_b = (int32_t)mp[P0.b];
        g.p0 = (int32_t)gpairs.size(); g.np = (int32_t)mem.size();
        for (int p : mem) {
            GPair q;
            q.hoff = ctx->pairs[p].hoff;
            cell_triplet(ctx, ctx->pair_ncn[p], q.ijk);
            q.pad = 0;
            gpairs.push_back(q);
        }
        for (int i0 = 0; i0 < g.noa; i0 += BL_TILE)
            for (int j0 = 0; j0 < g.nob; j0 += BL_TILE) {
                g.i0 = i0; g.j0 = j0;
                g.ni = std::min(BL_TILE, g.noa - i0); g.nj = std::min(BL_TILE, g.nob - j0);
                groups.push_back(g);
            }
        max_nel = std::max(max_nel, g.noa * g.nob);
    }
    BlochPlan *b = new BlochPlan();
    b->fsize = mp[ctx->natom];
    b->ngroups = (int32_t)groups.size();
    b->natom_pairs = (int64_t)members.size();
    b->max_nel = max_nel;
    if (hipMalloc(&b->d_groups, sizeof(GroupDev) * std::max<size_t>(1, groups.size())) != hipSuccess ||
        hipMalloc(&b->d_gpairs, sizeof(GPair) * std::max<size_t>(1, gpairs.size())) != hipSuccess ||
        (!groups.empty() && hipMemcpy(b->d_groups, groups.data(), sizeof(GroupDev) * groups.size(), hipMemcpyHostToDevice) != hipSuccess) ||
        (!gpairs.empty() && hipMemcpy(b->d_gpairs, gpairs.data(), sizeof(GPair) * gpairs.size(), hipMemcpyHostToDevice) != hipSuccess)) {
        bloch_free(b);
        return ctx_error(ctx, LCPAO_EDEVICE, "device allocation for the Bloch sums failed");
    }
    ctx->bloch = b;
    *out = b;
    return LCPAO_OK;
}

int bloch_stage(lcpao_ctx *ctx, BlochPlan *b, int64_t ndoubles) {
    if (ndoubles <= b->stage_n) return LCPAO_OK;
    if (b->stage) (void)hipFree(b->stage);
    b->stage = nullptr;
    b->stage_n = 0;
    HIP_TRY(ctx, hipMalloc(&b->stage, sizeof(double) * (size_t)ndoubles));
    b->stage_n = ndoubles;
    return LCPAO_OK;
}

// Zeroes the dense outputs (ld x ld complex each, a fill kernel) and runs
// k_bloch once for the nmat blocks described by A (flat, iflat, offsets,
// dense pointers filled in by the caller).
int bloch_launch(lcpao_ctx *ctx, BlochPlan *b, const double *kpt, BlochArgs &A, int nzero, double2 *const *zero) {
    for (int x = 0; x < 3; x++) A.k[x] = kpt[x];
    const int64_t nd = A.ld * A.ld;
    for (int z = 0; z < nzero; z++)
        if (nd > 0)
            hipLaunchKernelGGL(k_zero_c, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, ctx->stream, nd, zero[z]);
    if (b->ngroups > 0) {
        KTimer kt(ctx, KID_BLOCH);
        hipLaunchKernelGGL(k_bloch, dim3((unsigned)b->ngroups), dim3(BL_THREADS), 0, ctx->stream, b->d_groups,
                           b->d_gpairs, A);
    }
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    collect_timings(ctx);
    return LCPAO_OK;
}

int bloch_run(lcpao_ctx *ctx, const double *kpt, int32_t nmat, const double *const *d_flat, double *const *d_dense) {
    if (!kpt || !d_flat || !d_dense || nmat < 1 || nmat > BL_MAXMAT)
        return ctx_error(ctx, LCPAO_EINVAL, "null argument or nmat not in 1..4");
    for (int m = 0; m < nmat; m++)
        if (!d_flat[m] || !d_dense[m]) return ctx_error(ctx, LCPAO_EINVAL, "null matrix pointer");
    BlochPlan *b = nullptr;
    int rc = bloch_plan(ctx, &b);
    if (rc) return rc;
    BlochArgs A{};
    A.ld = b->fsize;
    A.nmat = nmat;
    double2 *zero[BL_MAXMAT];
    for (int m = 0; m < nmat; m++) {
        A.flat[m] = d_flat[m];
        A.dense[m] = zero[m] = reinterpret_cast<double2 *>(d_dense[m]);
    }
    return bloch_launch(ctx, b, kpt, A, nmat, zero);
}

// HS_matrix_NC! with k (HS_matrix.jl:211-257): one (2 fsize)^2 matrix, the
// up-up block from Hks[0] + im iHks[0], down-down from Hks[1] + im iHks[1],
// up-down (rows A up, columns B down) from Hks[2] + im (Hks[3] + iHks[2]);
// the down-up block stays zero, as in the reference.
int bloch_run_nc(lcpao_ctx *ctx, const double *kpt, const double *const *d_hks, const double *const *d_ihks,
                 double *d_H) {
    if (!kpt || !d_hks || !d_ihks || !d_H) return ctx_error(ctx, LCPAO_EINVAL, "null argument");
    for (int m = 0; m < 4; m++)
        if (!d_hks[m] || (m < 3 && !d_ihks[m])) return ctx_error(ctx, LCPAO_EINVAL, "null matrix pointer");
    BlochPlan *b = nullptr;
    int rc = bloch_plan(ctx, &b);
    if (rc) return rc;
    BlochArgs A{};
    A.ld = 2 * b->fsize;
    A.nmat = 3;
    double2 *H = reinterpret_cast<double2 *>(d_H);
    const int64_t f = b->fsize;
    const int64_t off_r[3] = {0, f, 0}, off_c[3] = {0, f, f};
    for (int m = 0; m < 3; m++) {
        A.flat[m] = d_hks[m];
        A.iflat[m][0] = m < 2 ? d_ihks[m] : d_hks[3];
        A.iflat[m][1] = m < 2 ? nullptr : d_ihks[2];
        A.off_r[m] = off_r[m];
        A.off_c[m] = off_c[m];
        A.dense[m] = H;
    }
    return bloch_launch(ctx, b, kpt, A, 1, &H);
}

// One spin block of the density matrix from states (see k_dm_states).
struct DmBlock {
    int64_t ra, rb;
    double *DM, *iDM;
};

int dm_run_blocks(lcpao_ctx *ctx, int32_t nk, const double *kpts, int64_t nstates, const double *const *d_C,
                  double scale, int64_t ld_factor, int nblk, const DmBlock *blk) {
    if (nk < 1 || !kpts || !d_C || nstates < 0) return ctx_error(ctx, LCPAO_EINVAL, "null argument or nk < 1");
    for (int k = 0; k < nk; k++)
        if (!d_C[k]) return ctx_error(ctx, LCPAO_EINVAL, "null state pointer");
    BlochPlan *b = nullptr;
    int rc = bloch_plan(ctx, &b);
    if (rc) return rc;
    // the k-point list and the state pointers travel in one device block
    const size_t bytes = sizeof(double) * 3 * nk + sizeof(void *) * nk;
    char *d_blk = nullptr;
    HIP_TRY(ctx, hipMalloc(&d_blk, bytes));
    std::vector<char> h(bytes);
    std::memcpy(h.data(), kpts, sizeof(double) * 3 * nk);
    std::memcpy(h.data() + sizeof(double) * 3 * nk, d_C, sizeof(void *) * nk);
    hipError_t e = hipMemcpy(d_blk, h.data(), bytes, hipMemcpyHostToDevice);
    for (int x = 0; x < nblk && e == hipSuccess && b->ngroups > 0; x++) {
        DmArgs A{};
        A.nk = nk;
        A.ld = ld_factor * b->fsize;
        A.nstates = nstates;
        A.ra = blk[x].ra;
        A.rb = blk[x].rb;
        A.scale = scale;
        A.kpts = reinterpret_cast<const double *>(d_blk);
        A.C = reinterpret_cast<const double2 *const *>(d_blk + sizeof(double) * 3 * nk);
        A.DM = blk[x].DM;
        A.iDM = blk[x].iDM;
        {
            KTimer kt(ctx, KID_DMK);
            hipLaunchKernelGGL(k_dm_states, dim3((unsigned)b->ngroups), dim3(BL_THREADS), 0, ctx->stream, b->d_groups,
                               b->d_gpairs, A);
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_blk);
    if (e != hipSuccess) return ctx_error(ctx, LCPAO_EDEVICE, std::string("density matrix: ") + hipGetErrorString(e));
    collect_timings(ctx);
    return LCPAO_OK;
}

int dm_run(lcpao_ctx *ctx, int32_t nk, const double *kpts, int64_t nstates, const double *const *d_C, double scale,
           double *d_DM) {
    if (!d_DM) return ctx_error(ctx, LCPAO_EINVAL, "null density-matrix pointer");
    const DmBlock blk{0, 0, d_DM, nullptr};
    return dm_run_blocks(ctx, nk, kpts, nstates, d_C, scale, 1, 1, &blk);
}

// Calc_DM_Crystal_NonCollinear! (Crystal_DFT.jl:342-410): DM[0..3] = Re uu,
// Re dd, Re ud, Im ud; with iDM (both or neither), the k-point sums of
// Calc_iDM_Crystal_NonCollinear! (Crystal_DFT.jl:413-486): iDM[0..1] =
// Im uu, Im dd, taken in the same passes.
int dm_run_nc(lcpao_ctx *ctx, int32_t nk, const double *kpts, int64_t nstates, const double *const *d_C, double scale,
              double *const *d_DM, double *const *d_iDM) {
    if (!d_DM || !d_DM[0] || !d_DM[1] || !d_DM[2] || !d_DM[3])
        return ctx_error(ctx, LCPAO_EINVAL, "null density-matrix pointer");
    if (d_iDM && (!d_iDM[0] || !d_iDM[1])) return ctx_error(ctx, LCPAO_EINVAL, "null iDM pointer");
    BlochPlan *b = nullptr;
    int rc = bloch_plan(ctx, &b);
    if (rc) return rc;
    const int64_t f = b->fsize;
    const DmBlock blk[3] = {{0, 0, d_DM[0], d_iDM ? d_iDM[0] : nullptr},
                            {f, f, d_DM[1], d_iDM ? d_iDM[1] : nullptr},
                            {0, f, d_DM[2], d_DM[3]}};
    return dm_run_blocks(ctx, nk, kpts, nstates, d_C, scale, 2, 3, blk);
}

}  // namespace
}  // namespace lcpao

extern "C" {

int lcpao_grid_bloch_sizes(lcpao_ctx *ctx, int64_t *fsize, int64_t *ngroups) {
    int rc = lcpao::ctx_enter(ctx);
    if (rc) return rc;
    lcpao::BlochPlan *b = nullptr;
    if ((rc = lcpao::bloch_plan(ctx, &b))) return rc;
    if (fsize) *fsize = b->fsize;
    if (ngroups) *ngroups = b->natom_pairs;
    return LCPAO_OK;
}

int lcpao_grid_bloch_sum_dev(lcpao_ctx *ctx, const double *kpt, int32_t nmat, const double *const *d_flat,
                             double *const *d_dense) {
    int rc = lcpao::ctx_enter(ctx);
    if (rc) return rc;
    return lcpao::bloch_run(ctx, kpt, nmat, d_flat, d_dense);
}

int lcpao_grid_bloch_sum(lcpao_ctx *ctx, const double *kpt, int32_t nmat, const double *const *flat,
                         double *const *dense) {
    int rc = lcpao::ctx_enter(ctx);
    if (rc) return rc;
    if (!flat || !dense || nmat < 1 || nmat > lcpao::BL_MAXMAT)
        return lcpao::ctx_error(ctx, LCPAO_EINVAL, "null argument or nmat not in 1..4");
    lcpao::BlochPlan *b = nullptr;
    if ((rc = lcpao::bloch_plan(ctx, &b))) return rc;
    const int64_t hs = std::max<int64_t>(1, ctx->my_hsize), nd = 2 * std::max<int64_t>(1, b->fsize * b->fsize);
    if ((rc = lcpao::bloch_stage(ctx, b, nmat * (hs + nd)))) return rc;
    const double *d_flat[lcpao::BL_MAXMAT];
    double *d_dense[lcpao::BL_MAXMAT];
    for (int m = 0; m < nmat; m++) {
        if (!flat[m] || !dense[m]) return lcpao::ctx_error(ctx, LCPAO_EINVAL, "null matrix pointer");
        d_flat[m] = b->stage + m * hs;
        d_dense[m] = b->stage + nmat * hs + m * nd;
        HIP_TRY(ctx, hipMemcpy((void *)d_flat[m], flat[m], sizeof(double) * ctx->my_hsize, hipMemcpyHostToDevice));
    }
    if ((rc = lcpao::bloch_run(ctx, kpt, nmat, d_flat, d_dense))) return rc;
    for (int m = 0; m < nmat; m++)
        HIP_TRY(ctx, hipMemcpy(dense[m], d_dense[m], sizeof(double) * 2 * b->fsize * b->fsize, hipMemcpyDeviceToHost));
    return LCPAO_OK;
}

int lcpao_grid_dm_states_dev(lcpao_ctx *ctx, int32_t nk, const double *kpts, int64_t nstates,
                             const double *const *d_C, double scale, double *d_DM) {
    int rc = lcpao::ctx_enter(ctx);
    if (rc) return rc;
    return lcpao::dm_run(ctx, nk, kpts, nstates, d_C, scale, d_DM);
}

int lcpao_grid_dm_states(lcpao_ctx *ctx, int32_t nk, const double *kpts, int64_t nstates, const double *const *C,
                         double scale, double *DM) {
    int rc = lcpao::ctx_enter(ctx);
    if (rc) return rc;
    if (nk < 1 || !C || !DM || nstates < 0) return lcpao::ctx_error(ctx, LCPAO_EINVAL, "null argument or nk < 1");
    lcpao::BlochPlan *b = nullptr;
    if ((rc = lcpao::bloch_plan(ctx, &b))) return rc;
    const int64_t per = 2 * std::max<int64_t>(1, b->fsize * nstates), hs = std::max<int64_t>(1, ctx->my_hsize);
    if ((rc = lcpao::bloch_stage(ctx, b, nk * per + hs))) return rc;
    std::vector<const double *> d_C(nk);
    for (int k = 0; k < nk; k++) {
        if (!C[k]) return lcpao::ctx_error(ctx, LCPAO_EINVAL, "null state pointer");
        d_C[k] = b->stage + k * per;
        HIP_TRY(ctx, hipMemcpy((void *)d_C[k], C[k], sizeof(double) * 2 * b->fsize * nstates, hipMemcpyHostToDevice));
    }
    double *d_DM = b->stage + nk * per;
    if ((rc = lcpao::dm_run(ctx, nk, kpts, nstates, d_C.data(), scale, d_DM))) return rc;
    HIP_TRY(ctx, hipMemcpy(DM, d_DM, sizeof(double) * ctx->my_hsize, hipMemcpyDeviceToHost));
    return LCPAO_OK;
}

int lcpao_grid_bloch_sum_nc_dev(lcpao_ctx *ctx, const double *kpt, const double *const *d_hks,
                                const double *const *d_ihks, double *d_H) {
    int rc = lcpao::ctx_enter(ctx);
    if (rc) return rc;
    return lcpao::bloch_run_nc(ctx, kpt, d_hks, d_ihks, d_H);
}

int lcpao_grid_bloch_sum_nc(lcpao_ctx *ctx, const double *kpt, const double *const *hks, const double *const *ihks,
                            double *H) {
    int rc = lcpao::ctx_enter(ctx);
    if (rc) return rc;
    if (!kpt || !hks || !ihks || !H) return lcpao::ctx_error(ctx, LCPAO_EINVAL, "null argument");
    for (int m = 0; m < 4; m++)
        if (!hks[m] || (m < 3 && !ihks[m])) return lcpao::ctx_error(ctx, LCPAO_EINVAL, "null matrix pointer");
    lcpao::BlochPlan *b = nullptr;
    if ((rc = lcpao::bloch_plan(ctx, &b))) return rc;
    const int64_t hs = std::max<int64_t>(1, ctx->my_hsize), nd = 2 * std::max<int64_t>(1, 4 * b->fsize * b->fsize);
    if ((rc = lcpao::bloch_stage(ctx, b, 7 * hs + nd))) return rc;
    const double *d_hks[4], *d_ihks[3];
    for (int m = 0; m < 7; m++) {
        const double *src = m < 4 ? hks[m] : ihks[m - 4];
        double *dst = b->stage + m * hs;
        HIP_TRY(ctx, hipMemcpy(dst, src, sizeof(double) * ctx->my_hsize, hipMemcpyHostToDevice));
        if (m < 4) d_hks[m] = dst;
        else d_ihks[m - 4] = dst;
    }
    double *d_H = b->stage + 7 * hs;
    if ((rc = lcpao::bloch_run_nc(ctx, kpt, d_hks, d_ihks, d_H))) return rc;
    HIP_TRY(ctx, hipMemcpy(H, d_H, sizeof(double) * 8 * b->fsize * b->fsize, hipMemcpyDeviceToHost));
    return LCPAO_OK;
}

int lcpao_grid_dm_states_nc_dev(lcpao_ctx *ctx, int32_t nk, const double *kpts, int64_t nstates,
                                const double *const *d_C, double scale, double *const *d_DM, double *const *d_iDM) {
    int rc = lcpao::ctx_enter(ctx);
    if (rc) return rc;
    return lcpao::dm_run_nc(ctx, nk, kpts, nstates, d_C, scale, d_DM, d_iDM);
}

int lcpao_grid_dm_states_nc(lcpao_ctx *ctx, int32_t nk, const double *kpts, int64_t nstates, const double *const *C,
                            double scale, double *const *DM, double *const *iDM) {
    int rc = lcpao::ctx_enter(ctx);
    if (rc) return rc;
    if (nk < 1 || !C || !DM || nstates < 0) return lcpao::ctx_error(ctx, LCPAO_EINVAL, "null argument or nk < 1");
    for (int m = 0; m < 4; m++)
        if (!DM[m]) return lcpao::ctx_error(ctx, LCPAO_EINVAL, "null density-matrix pointer");
    if (iDM && (!iDM[0] || !iDM[1])) return lcpao::ctx_error(ctx, LCPAO_EINVAL, "null iDM pointer");
    lcpao::BlochPlan *b = nullptr;
    if ((rc = lcpao::bloch_plan(ctx, &b))) return rc;
    const int64_t per = 4 * std::max<int64_t>(1, b->fsize * nstates), hs = std::max<int64_t>(1, ctx->my_hsize);
    const int nout = iDM ? 6 : 4;
    if ((rc = lcpao::bloch_stage(ctx, b, nk * per + nout * hs))) return rc;
    std::vector<const double *> d_C(nk);
    for (int k = 0; k < nk; k++) {
        if (!C[k]) return lcpao::ctx_error(ctx, LCPAO_EINVAL, "null state pointer");
        d_C[k] = b->stage + k * per;
        HIP_TRY(ctx, hipMemcpy((void *)d_C[k], C[k], sizeof(double) * 4 * b->fsize * nstates, hipMemcpyHostToDevice));
    }
    double *d_out[6];
    for (int m = 0; m < nout; m++) d_out[m] = b->stage + nk * per + m * hs;
    if ((rc = lcpao::dm_run_nc(ctx, nk, kpts, nstates, d_C.data(), scale, d_out, iDM ? d_out + 4 : nullptr)))
        return rc;
    for (int m = 0; m < nout; m++)
        HIP_TRY(ctx, hipMemcpy(m < 4 ? DM[m] : iDM[m - 4], d_out[m], sizeof(double) * ctx->my_hsize,
                               hipMemcpyDeviceToHost));
    return LCPAO_OK;
}

}  // extern "C"
