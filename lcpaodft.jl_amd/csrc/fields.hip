// fields.hip -- grid fields on either side of the contractions: the atomic
// and partial-core densities on the grid (Set_AdenPCC_Grid) and the Hartree
// potential of a density by FFT (Solve_Poisson!), both on the device so that
// the density and the potential stay in HBM between SCF steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "ctx.hpp"

namespace lcpao {

// ---------------------------------------------------------------------------
// Atomic densities
// ---------------------------------------------------------------------------
struct RadialDev {
    int32_t npao, nvps;
    const double *pxv, *prv, *pden;  // PAO log mesh, Spe_Atomic_Den (npao + 4)
    const double *vxv, *vrv, *vpcc;  // VPS log mesh, Spe_Atomic_PCC (nvps + 4)
};

// per atom: occupation shares and the spin direction's trigonometric factors
struct AtomSpinDev { double ocu, ocd, st, ct, cp, sp; };

struct FieldGeom {
    double glat[9], origin[3];
    int32_t n2, n3, natom, nspin;
};

// KumoF (utils/KumoF.jl:1-62): value at x = log r of a function tabulated on
// the log mesh (xv, rv); yv holds mesh point k at index k + 1.  Below the
// mesh a parabola a r^2 + b matched in value and slope to the Hermite cubic
// of the first interval pair; inside, the cubic through four neighbouring
// values.  Operation order as written there (no contraction).
__device__ double kumo(int N, double x, const double *__restrict__ xv, const double *__restrict__ rv,
                       const double *__restrict__ yv) {
    if (x < xv[0]) {
        const double r = exp(x);
        const double rm = rv[4];
        const double h1 = rv[3] - rv[2], h2 = rv[4] - rv[3], h3 = rv[5] - rv[4];
        const double f1 = yv[2], f2 = yv[3], f3 = yv[4], f4 = yv[5];
        const double g1 = ((f3 - f2) * h1 / h2 + (f2 - f1) * h2 / h1) / (h1 + h2);
        const double g2 = ((f4 - f3) * h2 / h3 + (f3 - f2) * h3 / h2) / (h2 + h3);
        const double x1 = rm - rv[3], x2 = rm - rv[4];
        const double y1 = x1 / h2, y2 = x2 / h2, y12 = y1 * y1, y22 = y2 * y2;
        const double A = 2 * f2 + h2 * g1, B = 2 * f3 - h2 * g2;
        const double C = 3 * f2 + h2 * g1, D = 3 * f3 - h2 * g2;
        const double E = C + A * y2, F = D - B * y1;
        const double f = y22 * E + y12 * F;
        const double df = (2 * y2 / h2) * E + y22 * A / h2 + (2 * y1 / h2) * F - y12 * B / h2;
        const double a = 0.5 * df / rm;
        const double b = f - a * rm * rm;
        return a * r * r + b;
    }
    const double xmin = xv[0], xmax = xv[N - 1];
    x = fmax(fmin(x, xmax), xmin);
    const double t = (double)(N - 1) * (x - xmin) / (xmax - xmin);
    const int i = (int)floor(t);
    const double dt = t - i;
    const double d0 = yv[i + 1] - yv[i], d1 = yv[i + 2] - yv[i + 1], d2 = yv[i + 3] - yv[i + 2];
    const double a = d0 - 2 * d1 + d2;
    const double b = -2 * d0 + 3 * d1 - d2;
    const double c = d0 + d1;
    return yv[i + 1] + 0.5 * (((a * dt + b) * dt + c) * dt);
}

// One thread per grid point g: the atoms' contributions at g in the order of
// the reference's atom loop (Set_AdenPCC_Grid.jl:56-115), read through the
// grid-point index of the atom lists (entries ascending = atoms ascending,
// then list order), so each sum is formed in the reference's order; the
// non-collinear rotation (:125-138) is applied at the end.
// (atom, cell) of every entry of the grid-point index, in its order: built
// once per context so that the atomic-density kernel reads one 8-byte record
// per list point instead of searching the owning atom.
__global__ void k_entry_info(int64_t nent, const int64_t *__restrict__ ent, const int64_t *__restrict__ gl_off,
                             int natom, const int32_t *__restrict__ clist, int2 *__restrict__ info) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nent) return;
    const int64_t e = ent[k];
    int lo = 0, hi = natom;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (gl_off[mid] <= e) lo = mid; else hi = mid;
    }
    info[k] = make_int2(lo, clist[e]);
}

__global__ void k_aden_pcc(int64_t nn, const int64_t *__restrict__ ent_off, const int2 *__restrict__ info,
                           const double *__restrict__ gxyz, const int32_t *__restrict__ a2s,
                           const double *__restrict__ atv, FieldGeom fg, const RadialDev *__restrict__ spe,
                           const AtomSpinDev *__restrict__ spin, double *__restrict__ aden, double *__restrict__ pcc,
                           double *__restrict__ rho0, double *__restrict__ rho1, double *__restrict__ rho2,
                           double *__restrict__ rho3) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nn) return;
    const int n1 = (int)(g / ((int64_t)fg.n2 * fg.n3));
    const int n2 = (int)((g - (int64_t)n1 * fg.n2 * fg.n3) / fg.n3);
    const int n3 = (int)(g - (int64_t)n1 * fg.n2 * fg.n3 - (int64_t)n2 * fg.n3);
    const double Cx = n1 * fg.glat[0] + n2 * fg.glat[3] + n3 * fg.glat[6] + fg.origin[0];
    const double Cy = n1 * fg.glat[1] + n2 * fg.glat[4] + n3 * fg.glat[7] + fg.origin[1];
    const double Cz = n1 * fg.glat[2] + n2 * fg.glat[5] + n3 * fg.glat[8] + fg.origin[2];
    double sa = 0.0, sp = 0.0, s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int64_t k = ent_off[g]; k < ent_off[g + 1]; k++) {
        const int2 ac = info[k];
        const int a = ac.x, c = ac.y;
        const double dx = Cx + atv[3 * c] - gxyz[3 * a];
        const double dy = Cy + atv[3 * c + 1] - gxyz[3 * a + 1];
        const double dz = Cz + atv[3 * c + 2] - gxyz[3 * a + 2];
        const double x = 0.5 * log(dx * dx + dy * dy + dz * dz);
        const RadialDev S = spe[a2s[a]];
        const double ad = kumo(S.npao, x, S.pxv, S.prv, S.pden);
        const double pd = kumo(S.nvps, x, S.vxv, S.vrv, S.vpcc);
        sa += 0.5 * ad;
        sp += 0.5 * pd;
        const AtomSpinDev P = spin[a];
        if (fg.nspin == 1) {
            s0 += 0.5 * ad;
        } else if (fg.nspin == 2) {
            s0 += P.ocu * ad;
            s1 += P.ocd * ad;
        } else {
            const double mag = (P.ocu - P.ocd) * ad;
            s0 += ad;
            s1 += mag * P.st * P.cp;
            s2 += mag * P.st * P.sp;
            s3 += mag * P.ct;
        }
    }
    aden[g] = sa;
    pcc[g] = sp;
    if (fg.nspin == 1) {
        rho0[g] = s0;
    } else if (fg.nspin == 2) {
        rho0[g] = s0;
        rho1[g] = s1;
    } else {
        rho0[g] = 0.5 * (s0 + s3);
        rho1[g] = 0.5 * (s0 - s3);
        rho2[g] = 0.5 * s1;
        rho3[g] = -0.5 * s2;
    }
}

// ---------------------------------------------------------------------------
// Poisson
// ---------------------------------------------------------------------------
// 3-D FFT of the NN complex values (n3 fastest) as three passes of 1-D
// transforms, one axis per pass.  A workgroup takes B lines of one axis into
// LDS (B consecutive lines along the fastest other index for the strided
// axes, so each element index is one coalesced run), runs the Stockham
// autosort stages there (mixed radix, twiddles from sincospi) and writes
// the lines back.  sign -1: FFTW's forward transform; +1: the unnormalised
// inverse.
constexpr int FFT_MAXST = 24;
constexpr int FFT_LDS_ELEMS = 4096;  // at most two buffers of 2048 complex values: 64 KiB
constexpr int FFT_THREADS = 256;

struct FftPass {
    int32_t n, nst, sign, B;
    int32_t rad[FFT_MAXST];
    int64_t es;       // element stride along the line
    int64_t tstride;  // stride between the B lines of a workgroup
    int64_t T;        // lines along t
    int64_t ostride;  // stride of the outer index
};

// radix-P butterfly i of a stage (Ns = product of the earlier radices)
template <int P>
__device__ __forceinline__ void fft_bfly(const double2 *__restrict__ src, double2 *__restrict__ dst, int n, int i,
                                         int Ns, int sign, const double2 *w) {
    const int m = n / P;
    const int j = i % Ns, k = i / Ns;
    double2 a[P];
#pragma unroll
    for (int r = 0; r < P; r++) {
        const double2 v = src[i + r * m];
        if (r == 0 || j == 0) {
            a[r] = v;
        } else {
            double sn, cs;
            sincospi(sign * 2.0 * (double)(r * j) / (double)(Ns * P), &sn, &cs);
            a[r] = make_double2(v.x * cs - v.y * sn, v.x * sn + v.y * cs);
        }
    }
    const int d = k * Ns * P + j;
#pragma unroll
    for (int q = 0; q < P; q++) {
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int r = 0; r < P; r++) {
            const double2 c = w[(r * q) % P];
            re += a[r].x * c.x - a[r].y * c.y;
            im += a[r].x * c.y + a[r].y * c.x;
        }
        dst[d + q * Ns] = make_double2(re, im);
    }
}

template <int P>
__device__ __forceinline__ void fft_stage(const double2 *src, double2 *dst, int n, int nb, int Ns, int sign) {
    double2 w[P];  // exp(sign 2 pi i t / P)
#pragma unroll
    for (int t = 0; t < P; t++) {
        double sn, cs;
        sincospi(sign * 2.0 * t / P, &sn, &cs);
        w[t] = make_double2(cs, sn);
    }
    const int m = n / P;
    for (int x = threadIdx.x; x < nb * m; x += blockDim.x) {
        const int b = x / m, i = x % m;
        fft_bfly<P>(src + b * n, dst + b * n, n, i, Ns, sign, w);
    }
}

// WIDE adds the radices 7, 11 and 13, whose unrolled butterflies need many
// more registers than 2 .. 5 (most grids factor into 2, 3 and 5 only).
template <bool WIDE>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_pass(double2 *__restrict__ data, FftPass P) {
    extern __shared__ double2 fft_lds[];
    const int n = P.n, B = P.B;
    const int64_t ngt = (P.T + B - 1) / B;
    const int64_t o = blockIdx.x / ngt;
    const int64_t t0 = (blockIdx.x % ngt) * B;
    const int nb = (int)(P.T - t0 < B ? P.T - t0 : B);
    const bool tfast = P.tstride == 1;
    double2 *src = fft_lds, *dst = fft_lds + B * n;
    const int64_t base = o * P.ostride + t0 * P.tstride;
    for (int x = threadIdx.x; x < nb * n; x += blockDim.x) {
        const int b = tfast ? x % nb : x / n, j = tfast ? x / nb : x % n;
        src[b * n + j] = data[base + b * P.tstride + j * P.es];
    }
    __syncthreads();
    int Ns = 1;
    for (int s = 0; s < P.nst; s++) {
        const int p = P.rad[s];
        switch (p) {
            case 2: fft_stage<2>(src, dst, n, nb, Ns, P.sign); break;
            case 3: fft_stage<3>(src, dst, n, nb, Ns, P.sign); break;
            case 4: fft_stage<4>(src, dst, n, nb, Ns, P.sign); break;
            case 5: fft_stage<5>(src, dst, n, nb, Ns, P.sign); break;
            default:
                if constexpr (WIDE) {
                    if (p == 7) fft_stage<7>(src, dst, n, nb, Ns, P.sign);
                    else if (p == 11) fft_stage<11>(src, dst, n, nb, Ns, P.sign);
                    else fft_stage<13>(src, dst, n, nb, Ns, P.sign);
                }
                break;
        }
        __syncthreads();
        double2 *t = src; src = dst; dst = t;
        Ns *= p;
    }
    for (int x = threadIdx.x; x < nb * n; x += blockDim.x) {
        const int b = tfast ? x % nb : x / n, j = tfast ? x / nb : x % n;
        data[base + b * P.tstride + j * P.es] = src[b * n + j];
    }
}

struct FieldPlan {
    int n[3] = {0, 0, 0};
    FftPass pass[3];          // per axis, sign set at launch
    bool wide[3] = {false, false, false};
    unsigned grid[3] = {0, 0, 0};
    bool ready = false;
    double2 *buf = nullptr;   // NN complex values
    double *stage = nullptr;  // host-pointer entry points: inputs and outputs
    int64_t stage_n = 0;
    int2 *entinfo = nullptr;  // (atom, cell) per grid-point index entry
};

void field_free(FieldPlan *f) {
    if (!f) return;
    if (f->buf) (void)hipFree(f->buf);
    if (f->stage) (void)hipFree(f->stage);
    if (f->entinfo) (void)hipFree(f->entinfo);
    delete f;
}

// Density_xyz of Set_Density_xyz! (Poisson.jl:10-26): nspin 1 -> 2 rho - 2 rho_atom;
// nspin 2 or 4 -> rho_up + rho_down - 2 rho_atom; nspin 0 -> rho as given
// (the inner solver, Poisson.jl:29).
__global__ void k_poisson_load(int64_t nn, int nspin, const double *__restrict__ r0, const double *__restrict__ r1,
                               const double *__restrict__ ra, double2 *__restrict__ buf) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nn) return;
    double v;
    if (nspin == 0) v = r0[g];
    else if (nspin == 1) v = 2 * r0[g] - 2 * ra[g];
    else v = r0[g] + r1[g] - 2 * ra[g];
    buf[g] = make_double2(v, 0.0);
}

// 4 pi rho(G) / |G|^2 with the reference's signed indices (Poisson.jl:65-80):
// index i maps to i below N/2 (integer division) and to i - N from there on,
// so for odd N the middle index counts as negative; G = 0 is set to zero.
// rec = Recvecs in column-major order (rec[3 j + i] = Recvecs[i+1, j+1]).
struct RecGeom { double rec[9]; int32_t n1, n2, n3, pad; };

__global__ void k_poisson_kernel(int64_t nn, RecGeom rg, double2 *__restrict__ buf) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nn) return;
    const int i = (int)(g / ((int64_t)rg.n2 * rg.n3));
    const int j = (int)((g / rg.n3) % rg.n2);
    const int k = (int)(g % rg.n3);
    const int sk1 = i < rg.n1 / 2 ? i : i - rg.n1;
    const int sk2 = j < rg.n2 / 2 ? j : j - rg.n2;
    const int sk3 = k < rg.n3 / 2 ? k : k - rg.n3;
    const double Gx = sk1 * rg.rec[0] + sk2 * rg.rec[1] + sk3 * rg.rec[2];
    const double Gy = sk1 * rg.rec[3] + sk2 * rg.rec[4] + sk3 * rg.rec[5];
    const double Gz = sk1 * rg.rec[6] + sk2 * rg.rec[7] + sk3 * rg.rec[8];
    const double temp = 4 * M_PI;
    const double g2 = Gx * Gx + Gy * Gy + Gz * Gz;
    double2 v = buf[g];
    if (g == 0) v = make_double2(0.0, 0.0);
    else v = make_double2(temp * v.x / g2, temp * v.y / g2);
    buf[g] = v;
}

// real part of the normalised inverse transform
__global__ void k_poisson_store(int64_t nn, const double2 *__restrict__ buf, double *__restrict__ v) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nn) return;
    v[g] = buf[g].x / (double)nn;
}

namespace {

// radices of n, fours first, then the primes 2 .. 13; 0 if n has a larger prime factor
int fft_radices(int n, int32_t *rad) {
    int k = 0;
    while (n % 4 == 0 && n > 4 && k < FFT_MAXST) { rad[k++] = 4; n /= 4; }
    const int primes[6] = {2, 3, 5, 7, 11, 13};
    for (int p : primes)
        while (n % p == 0 && k < FFT_MAXST) { rad[k++] = p; n /= p; }
    if (n == 4 && k < FFT_MAXST) { rad[k++] = 4; n = 1; }
    return n == 1 ? k : -1;
}

FieldPlan *fields_of(lcpao_ctx *ctx) {
    if (!ctx->fields) ctx->fields = new FieldPlan();
    return ctx->fields;
}

// The grid of a context never changes: the passes are set up once.
int field_plan(lcpao_ctx *ctx, FieldPlan **out) {
    FieldPlan *f = fields_of(ctx);
    if (f->ready) {
        *out = f;
        return LCPAO_OK;
    }
    const int64_t N1 = ctx->ngrid[0], N2 = ctx->ngrid[1], N3 = ctx->ngrid[2];
    for (int x = 0; x < 3; x++) {
        f->n[x] = ctx->ngrid[x];
        if (f->n[x] > 1024) return ctx_error(ctx, LCPAO_ERANGE, "grid axis longer than 1024 points");
        FftPass &P = f->pass[x];
        P = FftPass{};
        P.n = f->n[x];
        P.nst = fft_radices(P.n, P.rad);
        if (P.nst < 0) return ctx_error(ctx, LCPAO_ERANGE, "grid axis with a prime factor above 13");
        for (int k = 0; k < P.nst; k++) f->wide[x] = f->wide[x] || P.rad[k] > 5;
        // contiguous lines: B * n <= 1024 (32 KiB of LDS); strided axes: B a
        // multiple of 8 so that each element index is read as whole 128-byte
        // lines (B * n <= 2048)
        if (x == 2) P.B = std::max(1, 1024 / P.n);
        else P.B = P.n <= 256 ? (FFT_LDS_ELEMS / 2 / P.n) & ~7 : std::max(1, FFT_LDS_ELEMS / 2 / P.n);
        int64_t outer;
        if (x == 2) {         // contiguous lines, one after the other
            P.es = 1; P.tstride = N3; P.T = N1 * N2; P.ostride = 0; outer = 1;
        } else if (x == 1) {  // lines (n1, n3): B consecutive n3 per workgroup
            P.es = N3; P.tstride = 1; P.T = N3; P.ostride = N2 * N3; outer = N1;
        } else {              // lines (n2, n3) flattened
            P.es = N2 * N3; P.tstride = 1; P.T = N2 * N3; P.ostride = 0; outer = 1;
        }
        const int64_t g = outer * ((P.T + P.B - 1) / P.B);
        if (g > 0x7fffffff) return ctx_error(ctx, LCPAO_ERANGE, "grid too large for the FFT launch");
        f->grid[x] = (unsigned)g;
    }
    if (!f->buf) HIP_TRY(ctx, hipMalloc(&f->buf, sizeof(double2) * (size_t)std::max<int64_t>(1, ctx->nn)));
    f->ready = true;
    *out = f;
    return LCPAO_OK;
}

hipError_t fft_axis(lcpao_ctx *ctx, FieldPlan *f, int axis, int sign) {
    FftPass P = f->pass[axis];
    P.sign = sign;
    const size_t lds = sizeof(double2) * 2 * (size_t)P.B * P.n;
    KTimer kt(ctx, KID_FFT);
    if (f->wide[axis])
        hipLaunchKernelGGL(k_fft_pass<true>, dim3(f->grid[axis]), dim3(FFT_THREADS), lds, ctx->stream, f->buf, P);
    else
        hipLaunchKernelGGL(k_fft_pass<false>, dim3(f->grid[axis]), dim3(FFT_THREADS), lds, ctx->stream, f->buf, P);
    return hipGetLastError();
}

int field_stage(lcpao_ctx *ctx, FieldPlan *f, int64_t n) {
    if (n <= f->stage_n) return LCPAO_OK;
    if (f->stage) (void)hipFree(f->stage);
    f->stage = nullptr;
    f->stage_n = 0;
    HIP_TRY(ctx, hipMalloc(&f->stage, sizeof(double) * (size_t)n));
    f->stage_n = n;
    return LCPAO_OK;
}

int poisson_run(lcpao_ctx *ctx, int nspin, const double *const *d_rho, const double *d_aden, const double *rec,
                double *d_v) {
    if (nspin != 0 && nspin != 1 && nspin != 2 && nspin != 4) return ctx_error(ctx, LCPAO_EINVAL, "nspin must be 0, 1, 2 or 4");
    if (!d_rho || !d_rho[0] || (nspin >= 2 && !d_rho[1]) || (nspin >= 1 && !d_aden) || !rec || !d_v)
        return ctx_error(ctx, LCPAO_EINVAL, "null argument");
    FieldPlan *f = nullptr;
    int rc = field_plan(ctx, &f);
    if (rc) return rc;
    const int64_t nn = ctx->nn;
    const unsigned nb = (unsigned)((nn + 255) / 256);
    RecGeom rg;
    for (int x = 0; x < 9; x++) rg.rec[x] = rec[x];
    rg.n1 = ctx->ngrid[0]; rg.n2 = ctx->ngrid[1]; rg.n3 = ctx->ngrid[2]; rg.pad = 0;
    {
        KTimer kt(ctx, KID_PGRID);
        hipLaunchKernelGGL(k_poisson_load, dim3(nb), dim3(256), 0, ctx->stream, nn, nspin, d_rho[0],
                           nspin >= 2 ? d_rho[1] : nullptr, d_aden, f->buf);
    }
    HIP_TRY(ctx, hipGetLastError());
    // forward along n3, n2, n1 and back along n1, n2, n3, as Poisson.jl:37-104
    for (int axis = 2; axis >= 0; axis--) HIP_TRY(ctx, fft_axis(ctx, f, axis, -1));
    {
        KTimer kt(ctx, KID_PGRID);
        hipLaunchKernelGGL(k_poisson_kernel, dim3(nb), dim3(256), 0, ctx->stream, nn, rg, f->buf);
    }
    HIP_TRY(ctx, hipGetLastError());
    for (int axis = 0; axis < 3; axis++) HIP_TRY(ctx, fft_axis(ctx, f, axis, 1));
    {
        KTimer kt(ctx, KID_PGRID);
        hipLaunchKernelGGL(k_poisson_store, dim3(nb), dim3(256), 0, ctx->stream, nn, f->buf, d_v);
    }
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    collect_timings(ctx);
    return LCPAO_OK;
}

int aden_run(lcpao_ctx *ctx, int32_t nspin, int32_t nspecies, const lcpao_atomic_radial *spe,
             const double *init_nspin, const double *init_angle, double *d_aden, double *d_pcc,
             double *const *d_rho) {
    if (!ctx->have_lattice || !ctx->d_clist) return ctx_error(ctx, LCPAO_ESTATE, "geometry lacks lattice, positions, species or cell lists");
    if (nspin != 1 && nspin != 2 && nspin != 4) return ctx_error(ctx, LCPAO_EINVAL, "nspin must be 1, 2 or 4");
    if (nspecies <= 0 || !spe || !init_nspin || (nspin == 4 && !init_angle) || !d_aden || !d_pcc || !d_rho)
        return ctx_error(ctx, LCPAO_EINVAL, "null argument");
    for (int s = 0; s < nspin; s++)
        if (!d_rho[s]) return ctx_error(ctx, LCPAO_EINVAL, "null spin pointer");
    for (int a = 0; a < ctx->natom; a++)
        if (ctx->atom2spe[a] < 0 || ctx->atom2spe[a] >= nspecies) return ctx_error(ctx, LCPAO_EINVAL, "atom2spe out of range");
    for (int s = 0; s < nspecies; s++) {
        const lcpao_atomic_radial &R = spe[s];
        if (R.nmesh_pao < 6 || R.nmesh_vps < 6 || !R.pao_xv || !R.pao_rv || !R.atomic_den || !R.vps_xv || !R.vps_rv ||
            !R.atomic_pcc)
            return ctx_error(ctx, LCPAO_EINVAL, "bad radial description");
    }
    // one device block: per species the six arrays, then positions, cells, species, spins
    std::vector<double> blob;
    std::vector<RadialDev> rd(nspecies);
    std::vector<size_t> at(nspecies * 6);
    auto put = [&](const double *p, size_t n) { const size_t o = blob.size(); blob.insert(blob.end(), p, p + n); return o; };
    for (int s = 0; s < nspecies; s++) {
        const lcpao_atomic_radial &R = spe[s];
        at[6 * s + 0] = put(R.pao_xv, R.nmesh_pao);
        at[6 * s + 1] = put(R.pao_rv, R.nmesh_pao);
        at[6 * s + 2] = put(R.atomic_den, R.nmesh_pao + 4);
        at[6 * s + 3] = put(R.vps_xv, R.nmesh_vps);
        at[6 * s + 4] = put(R.vps_rv, R.nmesh_vps);
        at[6 * s + 5] = put(R.atomic_pcc, R.nmesh_vps + 4);
    }
    const size_t o_gxyz = put(ctx->gxyz.data(), ctx->gxyz.size());
    const std::vector<double> atv = cell_vectors(ctx);
    const size_t o_atv = put(atv.data(), atv.size());
    std::vector<AtomSpinDev> sp(ctx->natom);
    for (int a = 0; a < ctx->natom; a++) {
        // Set_AdenPCC_Grid.jl:70-78
        const double nu = init_nspin[2 * a], nd = init_nspin[2 * a + 1], ne = nu + nd;
        sp[a].ocu = ne > 1e-15 ? nu / ne : 0.0;
        sp[a].ocd = ne > 1e-15 ? nd / ne : 0.0;
        const double th = init_angle ? init_angle[2 * a] : 0.0, ph = init_angle ? init_angle[2 * a + 1] : 0.0;
        sp[a].st = std::sin(th); sp[a].ct = std::cos(th); sp[a].cp = std::cos(ph); sp[a].sp = std::sin(ph);
    }
    double *d_blob = nullptr;
    RadialDev *d_rd = nullptr;
    AtomSpinDev *d_sp = nullptr;
    int32_t *d_a2s = nullptr;
    bool ok = hipMalloc(&d_blob, sizeof(double) * blob.size()) == hipSuccess &&
              hipMalloc(&d_rd, sizeof(RadialDev) * nspecies) == hipSuccess &&
              hipMalloc(&d_sp, sizeof(AtomSpinDev) * ctx->natom) == hipSuccess &&
              hipMalloc(&d_a2s, sizeof(int32_t) * ctx->natom) == hipSuccess;
    int result = LCPAO_OK;
    if (ok) {
        for (int s = 0; s < nspecies; s++) {
            rd[s].npao = spe[s].nmesh_pao;
            rd[s].nvps = spe[s].nmesh_vps;
            rd[s].pxv = d_blob + at[6 * s + 0]; rd[s].prv = d_blob + at[6 * s + 1]; rd[s].pden = d_blob + at[6 * s + 2];
            rd[s].vxv = d_blob + at[6 * s + 3]; rd[s].vrv = d_blob + at[6 * s + 4]; rd[s].vpcc = d_blob + at[6 * s + 5];
        }
        ok = hipMemcpy(d_blob, blob.data(), sizeof(double) * blob.size(), hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(d_rd, rd.data(), sizeof(RadialDev) * nspecies, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(d_sp, sp.data(), sizeof(AtomSpinDev) * ctx->natom, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(d_a2s, ctx->atom2spe.data(), sizeof(int32_t) * ctx->natom, hipMemcpyHostToDevice) == hipSuccess;
    }
    if (!ok) {
        result = ctx_error(ctx, LCPAO_EDEVICE, "device allocation or upload for the atomic densities failed");
    } else if (ctx->nn > 0) {
        FieldGeom fg;
        for (int j = 0; j < 3; j++) {
            fg.glat[j] = ctx->lat[j] / ctx->ngrid[0];
            fg.glat[3 + j] = ctx->lat[3 + j] / ctx->ngrid[1];
            fg.glat[6 + j] = ctx->lat[6 + j] / ctx->ngrid[2];
            fg.origin[j] = ctx->origin[j];
        }
        fg.n2 = ctx->ngrid[1]; fg.n3 = ctx->ngrid[2]; fg.natom = ctx->natom; fg.nspin = nspin;
        const unsigned nb = (unsigned)((ctx->nn + 127) / 128);
        FieldPlan *f = fields_of(ctx);
        hipError_t e = hipSuccess;
        if (!f->entinfo) {
            const int64_t nent = std::max<int64_t>(1, ctx->total_gridn);
            e = hipMalloc(&f->entinfo, sizeof(int2) * (size_t)nent);
            if (e == hipSuccess && ctx->total_gridn > 0) {
                hipLaunchKernelGGL(k_entry_info, dim3((unsigned)((ctx->total_gridn + 255) / 256)), dim3(256), 0,
                                   ctx->stream, ctx->total_gridn, ctx->d_ent, ctx->d_gl_off, ctx->natom, ctx->d_clist,
                                   f->entinfo);
                e = hipGetLastError();
            }
            if (e != hipSuccess && f->entinfo) { (void)hipFree(f->entinfo); f->entinfo = nullptr; }
        }
        if (e == hipSuccess) {
            KTimer kt(ctx, KID_ADEN);
            hipLaunchKernelGGL(k_aden_pcc, dim3(nb), dim3(128), 0, ctx->stream, ctx->nn, ctx->d_ent_off, f->entinfo,
                               d_blob + o_gxyz, d_a2s, d_blob + o_atv, fg, d_rd, d_sp, d_aden, d_pcc, d_rho[0],
                               nspin >= 2 ? d_rho[1] : nullptr, nspin == 4 ? d_rho[2] : nullptr,
                               nspin == 4 ? d_rho[3] : nullptr);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) result = ctx_error(ctx, LCPAO_EDEVICE, std::string("atomic densities: ") + hipGetErrorString(e));
        else collect_timings(ctx);
    }
    (void)hipFree(d_blob); (void)hipFree(d_rd); (void)hipFree(d_sp); (void)hipFree(d_a2s);
    return result;
}

}  // namespace
}  // namespace lcpao

extern "C" {

int lcpao_grid_aden_pcc_dev(lcpao_ctx *ctx, int32_t nspin, int32_t nspecies, const lcpao_atomic_radial *spe,
                            const double *init_nspin, const double *init_angle, double *d_aden, double *d_pcc,
                            double *const *d_rho) {
    int rc = lcpao::ctx_enter(ctx);
    if (rc) return rc;
    return lcpao::aden_run(ctx, nspin, nspecies, spe, init_nspin, init_angle, d_aden, d_pcc, d_rho);
}

int lcpao_grid_aden_pcc(lcpao_ctx *ctx, int32_t nspin, int32_t nspecies, const lcpao_atomic_radial *spe,
                        const double *init_nspin, const double *init_angle, double *aden, double *pcc,
                        double *const *rho) {
    int rc = lcpao::ctx_enter(ctx);
    if (rc) return rc;
    if (!aden || !pcc || !rho || (nspin != 1 && nspin != 2 && nspin != 4))
        return lcpao::ctx_error(ctx, LCPAO_EINVAL, "null argument or nspin not 1, 2 or 4");
    for (int s = 0; s < nspin; s++)
        if (!rho[s]) return lcpao::ctx_error(ctx, LCPAO_EINVAL, "null spin pointer");
    const int64_t nn = ctx->nn;
    lcpao::FieldPlan *f = lcpao::fields_of(ctx);
    if ((rc = lcpao::field_stage(ctx, f, (2 + nspin) * std::max<int64_t>(1, nn)))) return rc;
    double *d_r[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int s = 0; s < nspin; s++) d_r[s] = f->stage + (2 + s) * nn;
    rc = lcpao::aden_run(ctx, nspin, nspecies, spe, init_nspin, init_angle, f->stage, f->stage + nn, d_r);
    if (rc) return rc;
    HIP_TRY(ctx, hipMemcpy(aden, f->stage, sizeof(double) * nn, hipMemcpyDeviceToHost));
    HIP_TRY(ctx, hipMemcpy(pcc, f->stage + nn, sizeof(double) * nn, hipMemcpyDeviceToHost));
    for (int s = 0; s < nspin; s++) HIP_TRY(ctx, hipMemcpy(rho[s], d_r[s], sizeof(double) * nn, hipMemcpyDeviceToHost));
    return LCPAO_OK;
}

int lcpao_grid_poisson_dev(lcpao_ctx *ctx, int32_t nspin, const double *const *d_rho, const double *d_aden,
                           const double *recvecs, double *d_vhart) {
    int rc = lcpao::ctx_enter(ctx);
    if (rc) return rc;
    return lcpao::poisson_run(ctx, nspin, d_rho, d_aden, recvecs, d_vhart);
}

int lcpao_grid_poisson(lcpao_ctx *ctx, int32_t nspin, const double *const *rho, const double *aden,
                       const double *recvecs, double *vhart) {
    int rc = lcpao::ctx_enter(ctx);
    if (rc) return rc;
    if (nspin != 0 && nspin != 1 && nspin != 2 && nspin != 4) return lcpao::ctx_error(ctx, LCPAO_EINVAL, "nspin must be 0, 1, 2 or 4");
    const int nin = nspin == 0 ? 1 : (nspin == 1 ? 1 : 2);
    if (!rho || !vhart || !recvecs || (nspin >= 1 && !aden)) return lcpao::ctx_error(ctx, LCPAO_EINVAL, "null argument");
    for (int s = 0; s < nin; s++)
        if (!rho[s]) return lcpao::ctx_error(ctx, LCPAO_EINVAL, "null spin pointer");
    const int64_t nn = ctx->nn;
    lcpao::FieldPlan *f = lcpao::fields_of(ctx);
    if ((rc = lcpao::field_stage(ctx, f, (nin + 2) * std::max<int64_t>(1, nn)))) return rc;
    const double *d_r[2] = {f->stage, f->stage + nn};
    for (int s = 0; s < nin; s++) HIP_TRY(ctx, hipMemcpy(f->stage + s * nn, rho[s], sizeof(double) * nn, hipMemcpyHostToDevice));
    double *d_a = f->stage + nin * nn, *d_v = f->stage + (nin + 1) * nn;
    if (nspin >= 1) HIP_TRY(ctx, hipMemcpy(d_a, aden, sizeof(double) * nn, hipMemcpyHostToDevice));
    rc = lcpao::poisson_run(ctx, nspin, d_r, nspin >= 1 ? d_a : nullptr, recvecs, d_v);
    if (rc) return rc;
    HIP_TRY(ctx, hipMemcpy(vhart, d_v, sizeof(double) * nn, hipMemcpyDeviceToHost));
    return LCPAO_OK;
}

}  // extern "C"
