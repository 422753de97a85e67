// ctx.hpp -- context of the grid library shared by its translation units:
// work-unit records, the lcpao_ctx structure, per-kernel event timing and
// the HIP error-check macro.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "lcpao_grid.h"

// overlap points per H / Force3 unit of the pair-list kernels: the units in
// flight (one wave each, launched in grid order) then touch few enough grid
// planes that the tables they gather stay in the Infinity Cache (GaAs216
// FP64 H 8.2 ms against 9.2 ms at 2048 points; Si511 FP64 Force3 44.6 ms
// against 54.8 ms).  H units write a partial block each (361 doubles for 19
// orbitals), so below 512 points their reduction costs more than the
// smaller band saves (256: 4.58 + 0.43 ms against 4.64 + 0.22 ms); Force3
// units write three sums and run at 256 (Si511 38.4 against 41.7 ms).
#define UNIT_PTS 256
#define HUNIT_PTS 512
#define RHO_CHUNK 2048    // grid-list entries per density unit (the largest decomposition)
#define HK 32             // points per LDS chunk in the H kernel
// orbital-count pairs (NO_A, NO_B) with kernels of fixed size (the Si/Cl,
// Na and Ga/As bases of the test systems); other counts take runtime loops
#define NO_CASES(X) X(13, 13) X(13, 14) X(14, 13) X(14, 14) X(19, 19) X(13, 19) X(19, 13) X(14, 19) X(19, 14)

// largest orbital count per atom: Read_PAO's basis string (Read_PAO.jl:49-76)
// takes one decimal digit of multiplicity for each of s, p, d, f, so at
// most 9 * (1 + 3 + 5 + 7) orbitals
#define MAXNO 144
#define MFMA64_NO 32      // the FP64 matrix-core density / Force3 kernels hold at most 32 orbitals
#define RHO_WAVES 4

namespace lcpao {

struct HUnit { int32_t pair, lo, n, pad; int64_t part; };
// Density work-unit decompositions: units of `chunk` consecutive grid-list
// entries of a home atom (2048, 1024 or 512), each holding the segments of
// the atom's computed pairs inside it.  The pair-list density kernels run on
// one of them with a number of waves per workgroup (autotuned).
constexpr int RHO_NDECOMP = 3;
constexpr int RHO_DECOMP_CHUNK[RHO_NDECOMP] = {2048, 1024, 512};
struct RhoUnit { int32_t atom, t0, nt, seg_beg, seg_end, pad; };
struct Seg { int32_t pair, lo, hi, pad; };
struct F3Unit { int32_t pair, lo, n, pad; };
struct PairDev { int32_t a, b, noa, nob; int64_t hoff, olg; int64_t tab_a, tab_b; int32_t gl_a, partner; };

struct SpeciesDev {
    int32_t maxl, nmesh, no, pad;
    int32_t num_basis[4];
    const double *rv;
    const double *rwf[4][16];   // multiplicity at most 16 per l
};

// Block-dense structure of the split-precision forms (built on first use).
struct BlockPlan;

// FFT plans and work buffers of the grid-field routines (Poisson solve),
// created on first use; released with the context.
struct FieldPlan;
void field_free(FieldPlan *f);

// Atom-pair groups of the Bloch sums (bloch.hip), built on first use.
struct BlochPlan;
void bloch_free(BlochPlan *b);

}  // namespace lcpao

using namespace lcpao;

// kernel ids of the per-kernel event timing (lcpao_grid_profile)
enum {
    KID_HAM = 0, KID_HRED = 1, KID_DEFF = 2, KID_RHO = 3, KID_RGATHER = 4, KID_F3 = 5, KID_F3RED = 6,
    KID_ADEN = 7, KID_FFT = 8, KID_PGRID = 9, KID_BLOCH = 10, KID_DMK = 11, KID_COUNT = 12
};

// tunable operations and their candidate counts
enum { TUNE_RHO64 = 0, TUNE_F364 = 1, TUNE_RHOBF = 2, TUNE_RHO32 = 3, TUNE_NOPS = 4 };
// FP64 density: 0 matrix cores, 1..5 vector FMA with unit shapes
// (waves x chunk) 4 x 2048, 4 x 1024, 8 x 512, 2 x 1024, 4 x 512, 6 the
// block form (vector FMA over the sub-blocks of the block plan, orbital
// columns staged in LDS), 7 the block form reading the columns from the
// tables; bf16x3 density: 0 single-buffered chunk planes, 1 double-buffered,
// 2 chunks loaded straight into LDS, 3 8-wave items (two workgroups per CU);
// FP32 density: 0 pair lists 4 x 2048, 1 block form, 2 pair lists 4 x 1024,
// 3 pair lists 8 x 512
constexpr int RHO64_BLOCK = 6, RHO64_BLOCK_TAB = 7, RHO64_BLOCK_MFMA = 8;
constexpr int TUNE_NCAND[TUNE_NOPS] = {9, 2, 4, 4};
constexpr int TUNE_MAXCAND = 9;

struct lcpao_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // geometry (host)
    int natom = 0, cpy = 0, maxno = 0;
    int ngrid[3] = {0, 0, 0};
    int64_t nn = 0, total_gridn = 0, total_olg = 0, npair = 0, my_hsize = 0;
    double gridvol = 0;
    std::vector<int> no, gridn;
    std::vector<int64_t> gl_off, tab_off;
    std::vector<PairDev> pairs;
    bool have_lattice = false;
    double lat[9], origin[3];
    std::vector<double> gxyz;
    std::vector<int> atom2spe;
    // device: lists and tables
    int32_t *d_glist = nullptr, *d_clist = nullptr;
    int64_t *d_gl_off = nullptr, *d_tab_off = nullptr;
    int32_t *d_no = nullptr;
    int32_t *d_olg1 = nullptr, *d_olg2 = nullptr;
    PairDev *d_pairs = nullptr;
    double *d_tab = nullptr, *d_dtab[3] = {nullptr, nullptr, nullptr};
    bool tables = false, dtables = false;
    // H
    std::vector<HUnit> hunits;
    HUnit *d_hunits = nullptr;
    int32_t *d_horder = nullptr;
    int32_t *d_hpair_ub = nullptr;  // canonical pair list + unit ranges
    int32_t *d_hcanon = nullptr;
    int64_t ncanon = 0, hpart_size = 0;
    int32_t horder_n = 0, forder_n = 0;  // entries of the XCD-dealt H / Force3 launch orders (-1: no unit)
    int32_t hred_threads = 64;   // FP64 H reduce width: the largest computed pair block, rounded to 64 (<= 1024)
    double *d_hpart = nullptr;
    // rho: the work-unit decompositions (RHO_DECOMP_CHUNK)
    struct RhoDecomp {
        int64_t nunits = 0;
        lcpao::RhoUnit *d_units = nullptr;
        int32_t *d_order = nullptr;
        lcpao::Seg *d_segs = nullptr;
    } rdec[lcpao::RHO_NDECOMP];
    double *d_deff = nullptr, *d_rhoA = nullptr;
    int64_t *d_ent_off = nullptr;
    int64_t *d_ent = nullptr;
    // Force3
    std::vector<F3Unit> funits;
    F3Unit *d_funits = nullptr;
    int32_t *d_forder = nullptr;
    int32_t *d_atom_ub = nullptr;
    double *d_fpart = nullptr, *d_fsum = nullptr;
    // staging for host-pointer entry points
    double *d_stage = nullptr;
    int64_t stage_size = 0;
    int32_t *d_flag = nullptr;
    // split-precision forms
    lcpao::BlockPlan *plan = nullptr;
    // block forms: -1 not decided yet, 1 the plan holds the geometry, 0 it
    // does not (too many orbital rows or sphere images on a grid block, or
    // more than 32 orbitals on an atom): the split forms then run the
    // pair-list kernels of a wider form (FP32 density and Force3, FP64 H)
    int block_state = -1;
    std::string block_why;
    // grid-field routines (atomic densities, Poisson)
    lcpao::FieldPlan *fields = nullptr;
    lcpao::BlochPlan *bloch = nullptr;
    float *d_tab32 = nullptr, *d_dtab32[3] = {nullptr, nullptr, nullptr};
    float *d_deff32 = nullptr;
    std::vector<int32_t> h_glist, h_clist;  // host copies for the block plan
    std::vector<int32_t> pair_ncn, numolg, canon;
    // per-kernel event timing
    bool own_stream = true;
    bool profiling = false;
    struct Pending { int id; hipEvent_t beg, end; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    // kernel variants chosen by the autotuner (-1: time the candidates on the
    // first call with this geometry), see lcpao_grid_tune
    int32_t variant[TUNE_NOPS] = {-1, -1, -1, -1};
    float variant_ms[TUNE_NOPS][TUNE_MAXCAND] = {};
    int64_t prof_launches[KID_COUNT] = {};
    double prof_ms[KID_COUNT] = {};
};




inline hipEvent_t take_event(lcpao_ctx *ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

// Brackets one kernel launch with events when profiling is on.
struct KTimer {
    lcpao_ctx *ctx;
    int id;
    hipEvent_t beg = nullptr;
    KTimer(lcpao_ctx *c, int i) : ctx(c), id(i) {
        if (ctx->profiling) {
            beg = take_event(ctx);
            (void)hipEventRecord(beg, ctx->stream);
        }
    }
    ~KTimer() {
        if (ctx->profiling && beg) {
            hipEvent_t end = take_event(ctx);
            (void)hipEventRecord(end, ctx->stream);
            ctx->pending.push_back({id, beg, end});
        }
    }
};

// After the stream has been synchronised: fold pending events into the counters.
inline void collect_timings(lcpao_ctx *ctx) {
    for (auto &p : ctx->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.beg, p.end) == hipSuccess) {
            ctx->prof_ms[p.id] += ms;
            ctx->prof_launches[p.id]++;
        }
        ctx->event_pool.push_back(p.beg);
        ctx->event_pool.push_back(p.end);
    }
    ctx->pending.clear();
}

#define HIP_TRY(ctx, call)                                                                     \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) {                                                                \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                   \
            return LCPAO_EDEVICE;                                                              \
        }                                                                                      \
    } while (0)

namespace lcpao {
// Shared host helpers of the translation units (lcpao_grid.hip).
int ctx_enter(lcpao_ctx *ctx);                          // null check + hipSetDevice
int ctx_error(lcpao_ctx *ctx, int code, const std::string &msg);
std::vector<double> cell_vectors(const lcpao_ctx *ctx);  // atv[cell][3], Generation_ATV order
void cell_triplet(const lcpao_ctx *ctx, int cell, int *ijk);  // atv_ijk[cell]
}  // namespace lcpao
