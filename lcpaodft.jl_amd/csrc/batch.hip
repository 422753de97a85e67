// batch.hip -- batched form: the split-BF16x3 grid Hamiltonian (1 channel)
// and density (1 spin) of many structures at once.  The block items of all
// structures are merged into one list that runs as one launch per
// operation, persistent with a work-queue counter or one workgroup per item
// (largest items first), so small structures no longer leave CUs idle and
// the per-structure launch and synchronisation gaps disappear.  The whole
// sequence (fold, fills, merged H launches, per-structure reductions, merged
// density launches) can be captured once in a HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "split.hpp"

using lcpao::BlockArgs;
using lcpao::HItem;

struct lcpao_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = true;
    std::string err;
    std::vector<lcpao_ctx *> ctx;
    struct Cls {
        int32_t rows, first, count;
    };
    std::vector<Cls> classes;
    HItem *d_items = nullptr;
    // density schedule 2: every structure's items of at most 8 super-tiles
    // (blocks of at most 384 rows) and the items of its larger blocks
    HItem *d_items8 = nullptr, *d_items_big = nullptr;
    int32_t nitems8 = 0, nitems_big = 0, max_rows_big = 0;
    BlockArgs *d_args_h = nullptr, *d_args_r = nullptr;
    int32_t *d_queue = nullptr;   // one counter per launch: 2 per class
    int max_slots = 0;
    // per-structure tables of the merged folds, fills and reductions
    lcpao::BatchFold *d_fold = nullptr;
    lcpao::BatchRed *d_red = nullptr;
    lcpao::BatchFill *d_fill = nullptr;
    int max_canon = 0, red_threads = 64, max_rows = 0;
    int32_t *d_canon_off = nullptr;  // prefix sums of the structures' computed pairs (flat fold / reduce grids)
    int32_t *d_red_perm = nullptr;   // the reduction's flat pairs, most contributions first
    int total_canon = 0;
    int64_t max_nn = 0;
    // LCPAO_BF16X3 (default) or LCPAO_FP32
    int32_t precision = LCPAO_BF16X3;
    // schedule of the bf16x3 launches: 0 persistent work queue, 1 one
    // workgroup per item, 2 as 1 with the density in items of at most 8
    // super-tiles, two workgroups per CU (-1: tune); the FP32 launches run
    // one workgroup per item
    int32_t sched = -1;
    float sched_ms[3] = {0.f, 0.f, 0.f};
    // arguments of the last call and the graph captured with them
    std::vector<const void *> bound;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

int berr(lcpao_batch *b, int code, const std::string &m) {
    b->err = m;
    return code;
}

#define BATCH_TRY(b, call)                                                                      \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) return berr(b, LCPAO_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

void drop_graph(lcpao_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Binds the per-structure device arguments of the merged launches.
int bind(lcpao_batch *b, const double *const *V, const double *const *DM, double *const *H, double *const *rho) {
    const size_t n = b->ctx.size();
    std::vector<const void *> now;
    for (size_t s = 0; s < n; s++) {
        if (!V[s] || !DM[s] || !H[s] || !rho[s]) return berr(b, LCPAO_EINVAL, "null structure pointer");
        now.insert(now.end(), {V[s], DM[s], H[s], rho[s]});
    }
    if (now == b->bound) return LCPAO_OK;
    std::vector<BlockArgs> ah(n), ar(n);
    std::vector<lcpao::BatchFold> fold(n);
    std::vector<lcpao::BatchRed> red(n);
    std::vector<lcpao::BatchFill> fill(n);
    for (size_t s = 0; s < n; s++) {
        lcpao_ctx *c = b->ctx[s];
        ah[s] = lcpao::block_args(c, V[s], nullptr);
        ar[s] = lcpao::block_args(c, nullptr, rho[s]);
        fold[s] = {c->d_hcanon, c->d_pairs, DM[s], c->d_deff32, (int32_t)c->ncanon, 0};
        red[s] = {c->plan->d_red_canon, c->plan->d_contrib_off, c->plan->d_contrib, c->d_pairs, c->plan->d_part, H[s],
                  c->gridvol, (int32_t)c->ncanon, 0};
        fill[s] = {rho[s], c->nn};
    }
    BATCH_TRY(b, hipMemcpy(b->d_args_h, ah.data(), n * sizeof(BlockArgs), hipMemcpyHostToDevice));
    BATCH_TRY(b, hipMemcpy(b->d_args_r, ar.data(), n * sizeof(BlockArgs), hipMemcpyHostToDevice));
    BATCH_TRY(b, hipMemcpy(b->d_fold, fold.data(), n * sizeof(lcpao::BatchFold), hipMemcpyHostToDevice));
    BATCH_TRY(b, hipMemcpy(b->d_red, red.data(), n * sizeof(lcpao::BatchRed), hipMemcpyHostToDevice));
    BATCH_TRY(b, hipMemcpy(b->d_fill, fill.data(), n * sizeof(lcpao::BatchFill), hipMemcpyHostToDevice));
    b->bound = now;
    drop_graph(b);  // the graph's small launches carry the old pointers
    return LCPAO_OK;
}

// Enqueues one batched H + density on b->stream (no synchronisation, no
// allocation, no memset: safe to capture): one launch each for the folds of
// every structure's density matrix (on-site blocks symmetrised), the cleared
// densities, the merged H items, the per-pair H reductions and the merged
// density items.
int enqueue(lcpao_batch *b, int sched) {
    const int n = (int)b->ctx.size();
    lcpao::launch_deff_f32_batch(b->stream, n, b->d_fold, b->d_canon_off, b->total_canon);
    lcpao::launch_fill_batch(b->stream, n, b->d_fill, b->max_nn);
    const BlockArgs none{};
    const bool f32 = b->precision == LCPAO_FP32;
    const bool queue = !f32 && sched == 0;
    if (queue) lcpao::launch_fill_i32(b->stream, 2 * (int64_t)b->classes.size(), b->d_queue, 0);
    for (size_t k = 0; k < b->classes.size(); k++) {
        const auto &cl = b->classes[k];
        if (f32)
            BATCH_TRY(b, lcpao::launch_f32_block_kernel(false, b->stream, cl.rows, b->max_slots, b->d_items + cl.first,
                                                        cl.count, b->d_args_h, none));
        else
            BATCH_TRY(b, lcpao::launch_block_kernel(false, b->stream, cl.rows, 0, b->d_items + cl.first, cl.count,
                                                    b->d_args_h, none, queue ? b->d_queue + 2 * k : nullptr,
                                                    queue ? 0u : (unsigned)cl.count));
    }
    lcpao::launch_hblk_reduce_batch(b->stream, n, b->d_red, b->d_canon_off, b->total_canon, b->red_threads,
                                    b->d_red_perm);
    if (!f32 && sched == 2) {
        BATCH_TRY(b, lcpao::launch_rblk_w8(b->stream, std::min(b->max_rows, 384), b->d_items8, b->nitems8, b->d_args_r,
                                           none));
        if (b->nitems_big > 0)
            BATCH_TRY(b, lcpao::launch_block_kernel(true, b->stream, b->max_rows_big, b->max_slots, b->d_items_big,
                                                    b->nitems_big, b->d_args_r, none, nullptr, (unsigned)b->nitems_big));
        BATCH_TRY(b, hipGetLastError());
        return LCPAO_OK;
    }
    for (size_t k = 0; k < b->classes.size(); k++) {
        const auto &cl = b->classes[k];
        if (f32)
            BATCH_TRY(b, lcpao::launch_f32_block_kernel(true, b->stream, cl.rows, b->max_slots, b->d_items + cl.first,
                                                        cl.count, b->d_args_r, none));
        else
            BATCH_TRY(b, lcpao::launch_block_kernel(true, b->stream, cl.rows, b->max_slots, b->d_items + cl.first,
                                                    cl.count, b->d_args_r, none, queue ? b->d_queue + 2 * k + 1 : nullptr,
                                                    queue ? 0u : (unsigned)cl.count));
    }
    BATCH_TRY(b, hipGetLastError());
    return LCPAO_OK;
}

}  // namespace

extern "C" {

int lcpao_batch_create(lcpao_batch **out, int32_t n, lcpao_ctx *const *ctxs) {
    if (!out) return LCPAO_EINVAL;
    *out = nullptr;
    if (n <= 0 || !ctxs) return LCPAO_EINVAL;
    auto *b = new lcpao_batch();
    int dev = -1;
    for (int s = 0; s < n; s++) {
        if (!ctxs[s]) { delete b; return LCPAO_EINVAL; }
        if (s == 0) dev = ctxs[s]->device;
        if (ctxs[s]->device != dev) { delete b; return LCPAO_EINVAL; }
        b->ctx.push_back(ctxs[s]);
    }
    b->device = dev;
    auto fail = [&](int rc) {
        lcpao_batch_destroy(b);
        return rc;
    };
    if (hipSetDevice(dev) != hipSuccess || hipStreamCreate(&b->stream) != hipSuccess) return fail(LCPAO_EDEVICE);
    // every structure's block plan and bf16 tiles; all items in one list,
    // largest (rows, then super-tiles) first, run as one launch per operation
    // (the kernels hold one workgroup per CU whatever R is)
    std::vector<HItem> items, items8, items_big;
    std::vector<int32_t> item_rows;
    int32_t max_rows = 0;
    for (int s = 0; s < n; s++) {
        lcpao_ctx *c = b->ctx[s];
        if (!c->tables) return fail(LCPAO_ESTATE);
        int rc = lcpao::split_tables(c, LCPAO_BF16X3);
        if (rc) return fail(rc);
        if (!c->d_deff) return fail(LCPAO_ESTATE);
        const lcpao::BlockPlan *P = c->plan;
        b->max_slots = std::max(b->max_slots, P->max_slots);
        max_rows = std::max(max_rows, P->max_rows);
        b->max_canon = std::max<int>(b->max_canon, (int)c->ncanon);
        b->red_threads = std::max(b->red_threads, P->red_threads);
        b->max_nn = std::max(b->max_nn, c->nn);
        for (HItem it : P->items) {
            it.pad = s;
            items.push_back(it);
            item_rows.push_back(P->blocks[it.block].rows);
        }
        // each structure's XCD-dealt order is kept (a block's items run back
        // to back on one XCD)
        for (HItem it : P->items8) {
            it.pad = s;
            items8.push_back(it);
        }
        for (HItem it : P->items_big) {
            it.pad = s;
            items_big.push_back(it);
        }
        b->max_rows_big = std::max(b->max_rows_big, P->max_rows_big);
    }
    std::vector<size_t> order(items.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
        return item_rows[x] != item_rows[y] ? item_rows[x] > item_rows[y] : items[x].nst > items[y].nst;
    });
    {
        std::vector<HItem> sorted(items.size());
        for (size_t i = 0; i < order.size(); i++) sorted[i] = items[order[i]];
        items.swap(sorted);
    }
    b->classes.push_back({max_rows, 0, (int32_t)items.size()});
    const size_t ni = std::max<size_t>(1, items.size());
    if (hipMalloc(&b->d_items, ni * sizeof(HItem)) != hipSuccess ||
        hipMalloc(&b->d_items8, std::max<size_t>(1, items8.size()) * sizeof(HItem)) != hipSuccess ||
        hipMalloc(&b->d_items_big, std::max<size_t>(1, items_big.size()) * sizeof(HItem)) != hipSuccess ||
        hipMalloc(&b->d_args_h, n * sizeof(BlockArgs)) != hipSuccess ||
        hipMalloc(&b->d_args_r, n * sizeof(BlockArgs)) != hipSuccess ||
        hipMalloc(&b->d_queue, std::max<size_t>(2, 2 * b->classes.size()) * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&b->d_fold, n * sizeof(lcpao::BatchFold)) != hipSuccess ||
        hipMalloc(&b->d_red, n * sizeof(lcpao::BatchRed)) != hipSuccess ||
        hipMalloc(&b->d_fill, n * sizeof(lcpao::BatchFill)) != hipSuccess ||
        hipMalloc(&b->d_canon_off, (n + 1) * sizeof(int32_t)) != hipSuccess)
        return fail(LCPAO_EDEVICE);
    {
        std::vector<int32_t> off(n + 1, 0);
        for (int s = 0; s < n; s++) {
            if ((int64_t)off[s] + b->ctx[s]->ncanon > INT32_MAX) return fail(LCPAO_ERANGE);
            off[s + 1] = off[s] + (int32_t)b->ctx[s]->ncanon;
        }
        b->total_canon = off[n];
        if (hipMemcpy(b->d_canon_off, off.data(), (n + 1) * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess)
            return fail(LCPAO_EDEVICE);
        // longest first over every structure's pairs (each structure's plan
        // already lists its pairs by falling contribution count)
        std::vector<int32_t> perm(b->total_canon), cnt(b->total_canon, 0);
        for (int s = 0; s < n; s++) {
            const auto &rc = b->ctx[s]->plan->red_count;
            for (int32_t k = 0; k < off[s + 1] - off[s] && k < (int32_t)rc.size(); k++) cnt[off[s] + k] = rc[k];
        }
        for (int32_t w = 0; w < b->total_canon; w++) perm[w] = w;
        std::stable_sort(perm.begin(), perm.end(), [&](int32_t x, int32_t y) { return cnt[x] > cnt[y]; });
        if (hipMalloc(&b->d_red_perm, std::max<size_t>(1, perm.size()) * sizeof(int32_t)) != hipSuccess ||
            (!perm.empty() &&
             hipMemcpy(b->d_red_perm, perm.data(), perm.size() * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess))
            return fail(LCPAO_EDEVICE);
    }
    b->max_rows = max_rows;
    b->nitems8 = (int32_t)items8.size();
    b->nitems_big = (int32_t)items_big.size();
    if (!items.empty() &&
        hipMemcpy(b->d_items, items.data(), items.size() * sizeof(HItem), hipMemcpyHostToDevice) != hipSuccess)
        return fail(LCPAO_EDEVICE);
    if (!items8.empty() &&
        hipMemcpy(b->d_items8, items8.data(), items8.size() * sizeof(HItem), hipMemcpyHostToDevice) != hipSuccess)
        return fail(LCPAO_EDEVICE);
    if (!items_big.empty() &&
        hipMemcpy(b->d_items_big, items_big.data(), items_big.size() * sizeof(HItem), hipMemcpyHostToDevice) != hipSuccess)
        return fail(LCPAO_EDEVICE);
    *out = b;
    return LCPAO_OK;
}

int lcpao_batch_assemble(lcpao_batch *b, const double *const *d_V, const double *const *d_DM, double *const *d_H,
                         double *const *d_rho, int32_t use_graph) {
    if (!b) return LCPAO_EINVAL;
    if (!d_V || !d_DM || !d_H || !d_rho) return berr(b, LCPAO_EINVAL, "null pointer array");
    BATCH_TRY(b, hipSetDevice(b->device));
    // a member whose tables changed since the last call (set_tables,
    // set_dtables, build_tables clear its tiles' ready flag) gets its bf16
    // tiles rebuilt into the same buffers, so the bound arguments and a
    // captured graph stay valid
    for (lcpao_ctx *c : b->ctx) {
        if (!c->tables) return berr(b, LCPAO_ESTATE, "a member context has no orbital tables");
        const bool ready = c->plan && (b->precision == LCPAO_FP32 ? c->plan->f32_ready : c->plan->bf16_ready);
        if (!ready) {
            const bool had_t32 = c->plan && c->plan->d_t32;
            const int rc = lcpao::split_tables(c, b->precision);
            if (rc) return berr(b, rc, "rebuilding a member's split tables: " + c->err);
            if (b->precision == LCPAO_FP32 && !had_t32) b->bound.clear();  // new tile buffer: bind again
        }
    }
    int rc = bind(b, d_V, d_DM, d_H, d_rho);
    if (rc) return rc;
    if (b->precision == LCPAO_FP32 && b->sched < 0) b->sched = 1;  // the FP32 launches: one workgroup per item
    if (b->sched < 0) {  // tune the schedule: each candidate once untimed, once timed
        hipEvent_t e0, e1;
        BATCH_TRY(b, hipEventCreate(&e0));
        BATCH_TRY(b, hipEventCreate(&e1));
        int best = 0;
        for (int v = 0; v < 3; v++) {
            if ((rc = enqueue(b, v))) return rc;
            BATCH_TRY(b, hipEventRecord(e0, b->stream));
            if ((rc = enqueue(b, v))) return rc;
            BATCH_TRY(b, hipEventRecord(e1, b->stream));
            BATCH_TRY(b, hipEventSynchronize(e1));
            BATCH_TRY(b, hipEventElapsedTime(&b->sched_ms[v], e0, e1));
            if (b->sched_ms[v] < b->sched_ms[best]) best = v;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        b->sched = best;
    }
    if (!use_graph) {
        if ((rc = enqueue(b, b->sched))) return rc;
    } else {
        if (!b->exec) {
            BATCH_TRY(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue(b, b->sched);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
            if (e != hipSuccess) return berr(b, LCPAO_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            b->graph = g;
            BATCH_TRY(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        }
        BATCH_TRY(b, hipGraphLaunch(b->exec, b->stream));
    }
    BATCH_TRY(b, hipStreamSynchronize(b->stream));
    return LCPAO_OK;
}

int lcpao_batch_set_precision(lcpao_batch *b, int32_t precision) {
    if (!b) return LCPAO_EINVAL;
    if (precision != LCPAO_BF16X3 && precision != LCPAO_FP32)
        return berr(b, LCPAO_EINVAL, "the batched form runs LCPAO_BF16X3 or LCPAO_FP32");
    if (precision != b->precision) {
        drop_graph(b);
        b->bound.clear();
        b->precision = precision;
        b->sched = precision == LCPAO_FP32 ? 1 : -1;
        b->sched_ms[0] = b->sched_ms[1] = b->sched_ms[2] = 0.f;
    }
    return LCPAO_OK;
}

int lcpao_batch_tune(lcpao_batch *b, int32_t sched) {
    if (!b || sched < -1 || sched > 2) return LCPAO_EINVAL;
    if (sched != b->sched) drop_graph(b);
    b->sched = sched;
    return LCPAO_OK;
}

int lcpao_batch_tuned(const lcpao_batch *b, int32_t *sched, float *candidate_ms) {
    if (!b || !sched) return LCPAO_EINVAL;
    *sched = b->sched;
    if (candidate_ms)
        for (int v = 0; v < 3; v++) candidate_ms[v] = b->sched_ms[v];
    return LCPAO_OK;
}

int lcpao_batch_sizes(const lcpao_batch *b, int32_t *nstruct, int32_t *nclasses, int64_t *nitems) {
    if (!b) return LCPAO_EINVAL;
    if (nstruct) *nstruct = (int32_t)b->ctx.size();
    if (nclasses) *nclasses = (int32_t)b->classes.size();
    if (nitems) {
        int64_t t = 0;
        for (const auto &c : b->classes) t += c.count;
        *nitems = t;
    }
    return LCPAO_OK;
}

const char *lcpao_batch_last_error(const lcpao_batch *b) { return b ? b->err.c_str() : ""; }

void lcpao_batch_destroy(lcpao_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    drop_graph(b);
    void *ptrs[] = {b->d_items, b->d_items8, b->d_items_big, b->d_args_h, b->d_args_r, b->d_queue, b->d_fold, b->d_red, b->d_fill,
                    b->d_canon_off, b->d_red_perm};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // extern "C"
