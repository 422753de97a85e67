// ucell.cpp -- host-side construction of the UCell lists that the grid
// contractions consume: neighbour pairs, per-atom grid lists, the rank-local
// pair split and the pair overlap lists.
//
// The integer outputs must equal the reference's exactly, and several of
// them depend on floating-point boundary decisions (R <= rcut), so the
// expressions that decide membership keep the reference's operation order:
//   Get_FNAN / Trn_System       UCell.jl:520-672
//   Generation_ATV!             Generation_ATV.jl:1-38
//   Calc_AtomsGrid              UCell.jl:297-411 (+ Find_CGrids!, Find_CGrid.jl:1-49)
//   split_system_grid           UCell.jl:50-132
// The overlap lists (Calc_AtomOverlap_Grid, UCell.jl:425-516) are built by a
// different algorithm with the same result: an inverse index of the sphere
// entries covering each home grid point, visited in grid order, yields every
// pair's list already ordered by the B-atom index as the reference's merge
// does.  The work is split over host threads.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "lcpao_grid.h"

namespace {

thread_local std::string g_ucell_err;

constexpr double kPi = 3.141592653589793;

struct UCell {
    int natom = 0, cpy = 0, ncell = 0, nr = 0;
    double lat[9];
    int ngrid[3];
    double origin[3];
    double gridvol = 0;
    std::vector<double> gxyz, cut;
    std::vector<int> no;
    std::vector<double> atv;
    std::vector<int> atv_ijk, ratv;
    std::vector<int> fnan, natn, ncn;
    std::vector<int64_t> nb_off;
    std::vector<int> gridn, glist, clist;
    std::vector<int64_t> gl_off;
    int nprocs = 1, rank = 0;
    int64_t nloop = 0, npair = 0, total_hsize = 0, my_hsize = 0, my_hoff = 0;
    std::vector<int> p_atom, p_slot, p_natn, p_ncn;
    std::vector<int64_t> p_hoff;
    std::vector<int> numolg, olg1, olg2;
    std::vector<int64_t> olg_off;
};

template <class F>
void parallel_for(int64_t n, int nthreads, F f) {
    if (nthreads <= 1 || n <= 1) {
        for (int64_t i = 0; i < n; i++) f(i, 0);
        return;
    }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++)
        th.emplace_back([&, t] {
            for (;;) {
                int64_t i = next.fetch_add(1);
                if (i >= n) break;
                f(i, t);
            }
        });
    for (auto &x : th) x.join();
}

// Lattice translations in Generation_ATV! order: cell 0 = home, then the
// (i,j,k) loop over [-N,N]^3 with (0,0,0) skipped.
std::vector<double> make_atv(const double *lat, int N) {
    int nc = (2 * N + 1) * (2 * N + 1) * (2 * N + 1);
    std::vector<double> atv((size_t)nc * 3, 0.0);
    int rn = 1;
    for (int i = -N; i <= N; i++)
        for (int j = -N; j <= N; j++)
            for (int k = -N; k <= N; k++) {
                if (i == 0 && j == 0 && k == 0) continue;
                const double di = i, dj = j, dk = k;
                double *a = &atv[3 * (size_t)(rn++)];
                a[0] = di * lat[0] + dj * lat[3] + dk * lat[6];
                a[1] = di * lat[1] + dj * lat[4] + dk * lat[7];
                a[2] = di * lat[2] + dj * lat[5] + dk * lat[8];
            }
    return atv;
}

// Neighbour count of atom a with translations atv (the distance test of
// Estimate_Trn_System / Trn_System); with out != nullptr, records them.
int count_neighbours(const UCell &U, int a, const std::vector<double> &atv, int *natn_out, int *ncn_out) {
    const int nc = (int)(atv.size() / 3);
    const double *g1 = &U.gxyz[3 * (size_t)a];
    int f = 0;
    if (natn_out) { natn_out[0] = a; ncn_out[0] = 0; }
    for (int b = 0; b < U.natom; b++) {
        const double *g2 = &U.gxyz[3 * (size_t)b];
        const double rcut = U.cut[a] + U.cut[b];
        for (int rn = 0; rn < nc; rn++) {
            if (a == b && rn == 0) continue;
            const double dx = std::fabs(g1[0] - g2[0] - atv[3 * (size_t)rn]);
            const double dy = std::fabs(g1[1] - g2[1] - atv[3 * (size_t)rn + 1]);
            const double dz = std::fabs(g1[2] - g2[2] - atv[3 * (size_t)rn + 2]);
            if (dx <= rcut && dy <= rcut && dz <= rcut) {
                const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
                if (r <= rcut) {
                    f++;
                    if (natn_out) { natn_out[f] = b; ncn_out[f] = rn; }
                }
            }
        }
    }
    return f;
}

bool neighbours(UCell &U, int nthreads) {
    // CpyCell grows until the neighbour total stops changing (UCell.jl:530-550).
    int po = 0, cpy = 0, count = 1;
    const int countmax = 7;
    int64_t tf = 0;
    std::vector<int> cnt(U.natom);
    while (po == 0 && count < countmax) {
        cpy++;
        auto atv = make_atv(U.lat, cpy);
        int64_t prev = tf;
        parallel_for(U.natom, nthreads, [&](int64_t a, int) { cnt[a] = count_neighbours(U, (int)a, atv, nullptr, nullptr); });
        tf = 0;
        for (int a = 0; a < U.natom; a++) tf += cnt[a];
        if (tf == prev) po = 1;
        count++;
    }
    if (count == countmax || po == 0) { g_ucell_err = "neighbour search did not converge (Get_FNAN)"; return false; }
    U.cpy = cpy;
    auto atv = make_atv(U.lat, cpy);
    parallel_for(U.natom, nthreads, [&](int64_t a, int) { cnt[a] = count_neighbours(U, (int)a, atv, nullptr, nullptr); });
    U.fnan = cnt;
    U.nb_off.assign(U.natom + 1, 0);
    for (int a = 0; a < U.natom; a++) U.nb_off[a + 1] = U.nb_off[a] + U.fnan[a] + 1;
    U.natn.assign(U.nb_off[U.natom], 0);
    U.ncn.assign(U.nb_off[U.natom], 0);
    parallel_for(U.natom, nthreads, [&](int64_t a, int) {
        count_neighbours(U, (int)a, atv, &U.natn[U.nb_off[a]], &U.ncn[U.nb_off[a]]);
    });
    return true;
}

void gen_cells(UCell &U) {
    const int N = U.cpy;
    U.ncell = (2 * N + 1) * (2 * N + 1) * (2 * N + 1);
    U.nr = 2 * N + 4;
    U.atv = make_atv(U.lat, N);
    U.atv_ijk.assign((size_t)U.ncell * 3, 0);
    U.ratv.assign((size_t)U.nr * U.nr * U.nr, 0);
    int rn = 1;
    for (int i = -N; i <= N; i++)
        for (int j = -N; j <= N; j++)
            for (int k = -N; k <= N; k++) {
                int slot = (i == 0 && j == 0 && k == 0) ? 0 : rn++;
                U.atv_ijk[3 * (size_t)slot] = i;
                U.atv_ijk[3 * (size_t)slot + 1] = j;
                U.atv_ijk[3 * (size_t)slot + 2] = k;
                U.ratv[((size_t)(i + N) * U.nr + (j + N)) * U.nr + (k + N)] = slot;
            }
}

// 3x3 inverse and determinant by LU with partial pivoting.
void inv3(const double *A, double *Ainv) {
    double a[3][3], inv[3][3];
    int perm[3] = {0, 1, 2};
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) a[i][j] = A[3 * i + j];
    for (int k = 0; k < 3; k++) {
        int p = k;
        for (int i = k + 1; i < 3; i++) if (std::fabs(a[i][k]) > std::fabs(a[p][k])) p = i;
        if (p != k) {
            for (int j = 0; j < 3; j++) std::swap(a[k][j], a[p][j]);
            std::swap(perm[k], perm[p]);
        }
        for (int i = k + 1; i < 3; i++) {
            a[i][k] /= a[k][k];
            for (int j = k + 1; j < 3; j++) a[i][j] -= a[i][k] * a[k][j];
        }
    }
    for (int c = 0; c < 3; c++) {
        double y[3];
        for (int i = 0; i < 3; i++) {
            double s = (perm[i] == c) ? 1.0 : 0.0;
            for (int j = 0; j < i; j++) s -= a[i][j] * y[j];
            y[i] = s;
        }
        for (int i = 2; i >= 0; i--) {
            double s = y[i];
            for (int j = i + 1; j < 3; j++) s -= a[i][j] * inv[j][c];
            inv[i][c] = s / a[i][i];
        }
    }
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) Ainv[3 * i + j] = inv[i][j];
}

double det3(const double *A) {
    double a[3][3], sgn = 1.0;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) a[i][j] = A[3 * i + j];
    for (int k = 0; k < 3; k++) {
        int p = k;
        for (int i = k + 1; i < 3; i++) if (std::fabs(a[i][k]) > std::fabs(a[p][k])) p = i;
        if (p != k) { for (int j = 0; j < 3; j++) std::swap(a[k][j], a[p][j]); sgn = -sgn; }
        for (int i = k + 1; i < 3; i++) {
            double l = a[i][k] / a[k][k];
            for (int j = k + 1; j < 3; j++) a[i][j] -= l * a[k][j];
        }
    }
    return sgn * (a[0][0] * a[1][1] * a[2][2]);
}

// Decomposes an unbounded grid coordinate into (image index l0, home index n1)
// with the reference's integer arithmetic (Find_ln, Find_CGrid.jl:31-49).
inline void find_ln(int N, int n0, int &l0, int &n1) {
    if (n0 < 0) {
        int an = -n0;
        l0 = -(int)((double)(an - 1) / (double)N) - 1;
        int al = l0 < 0 ? -l0 : l0;
        n1 = N - (an - N * (al - 1));
    } else if (N <= n0) {
        l0 = (int)((double)n0 / (double)N);
        n1 = n0 - N * l0;
    } else {
        l0 = 0;
        n1 = n0;
    }
}

bool atoms_grid(UCell &U, int nthreads) {
    const int N1 = U.ngrid[0], N2 = U.ngrid[1], N3 = U.ngrid[2];
    double glat[9];
    for (int j = 0; j < 3; j++) {
        glat[j] = U.lat[j] / N1;
        glat[3 + j] = U.lat[3 + j] / N2;
        glat[6 + j] = U.lat[6 + j] / N3;
    }
    double glt[9], ginv[9], grec[9];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) glt[3 * i + j] = glat[3 * j + i];
    inv3(glt, ginv);
    for (int i = 0; i < 9; i++) grec[i] = 2 * kPi * ginv[i];
    static const int kidx[3][2] = {{1, 2}, {2, 0}, {0, 1}};
    struct Entry { int g, c; };
    std::vector<std::vector<Entry>> lists(U.natom);
    // The reference keeps the last computed point position when a box point
    // falls outside the CpyCell range (Find_CGrid.jl:140-144); that carries a
    // value across atoms, so the scan runs serially when it can happen.
    std::atomic<bool> stale_case{false};
    std::atomic<bool> bounds_error{false};
    auto scan = [&](int a, double *Cxyz) {
        const double Gx = U.gxyz[3 * (size_t)a], Gy = U.gxyz[3 * (size_t)a + 1], Gz = U.gxyz[3 * (size_t)a + 2];
        double rcut = U.cut[a] + 0.5;
        int nmin[3], nmax[3];
        for (int k = 0; k < 3; k++) {
            const double *Li = U.lat + 3 * kidx[k][0], *Lj = U.lat + 3 * kidx[k][1];
            double vx = Li[1] * Lj[2] - Li[2] * Lj[1];
            double vy = Li[2] * Lj[0] - Li[0] * Lj[2];
            double vz = Li[0] * Lj[1] - Li[1] * Lj[0];
            const double coef = 1.0 / std::sqrt(vx * vx + vy * vy + vz * vz);
            vx = vx * coef; vy = vy * coef; vz = vz * coef;
            double Cx = Gx + rcut * vx - U.origin[0], Cy = Gy + rcut * vy - U.origin[1], Cz = Gz + rcut * vz - U.origin[2];
            nmax[k] = (int)((Cx * grec[3 * k] + Cy * grec[3 * k + 1] + Cz * grec[3 * k + 2]) * 0.5 / kPi);
            Cx = Gx - rcut * vx - U.origin[0]; Cy = Gy - rcut * vy - U.origin[1]; Cz = Gz - rcut * vz - U.origin[2];
            nmin[k] = (int)((Cx * grec[3 * k] + Cy * grec[3 * k + 1] + Cz * grec[3 * k + 2]) * 0.5 / kPi);
            if (nmax[k] < nmin[k]) std::swap(nmin[k], nmax[k]);
        }
        rcut = U.cut[a];
        auto &L = lists[a];
        L.clear();
        for (int n1 = nmin[0]; n1 <= nmax[0]; n1++)
            for (int n2 = nmin[1]; n2 <= nmax[1]; n2++)
                for (int n3 = nmin[2]; n3 <= nmax[2]; n3++) {
                    int l1, l2, l3, m1, m2, m3;
                    find_ln(N1, n1, l1, m1);
                    find_ln(N2, n2, l2, m2);
                    find_ln(N3, n3, l3, m3);
                    int rn;
                    if (U.cpy < std::abs(l1 + 1) || U.cpy < std::abs(l2 + 1) || U.cpy < std::abs(l3 + 1)) {
                        rn = 0;
                        stale_case = true;
                    } else {
                        const int i1 = U.cpy + l1, i2 = U.cpy + l2, i3 = U.cpy + l3;
                        if (i1 < 0 || i2 < 0 || i3 < 0) { bounds_error = true; return; }
                        rn = U.ratv[((size_t)i1 * U.nr + i2) * U.nr + i3];
                        const double *av = &U.atv[3 * (size_t)rn];
                        Cxyz[0] = av[0] + m1 * glat[0] + m2 * glat[3] + m3 * glat[6] + U.origin[0];
                        Cxyz[1] = av[1] + m1 * glat[1] + m2 * glat[4] + m3 * glat[7] + U.origin[1];
                        Cxyz[2] = av[2] + m1 * glat[2] + m2 * glat[5] + m3 * glat[8] + U.origin[2];
                    }
                    const int g = m1 * N2 * N3 + m2 * N3 + m3;
                    const double dx = Cxyz[0] - Gx, dy = Cxyz[1] - Gy, dz = Cxyz[2] - Gz;
                    const double R = std::sqrt(dx * dx + dy * dy + dz * dz);
                    if (R <= rcut) L.push_back({g, rn});
                }
        std::stable_sort(L.begin(), L.end(), [](const Entry &x, const Entry &y) { return x.g < y.g; });
    };
    parallel_for(U.natom, nthreads, [&](int64_t a, int) {
        double C0[3] = {0.0, 0.0, 0.0};
        scan((int)a, C0);
    });
    if (stale_case) {  // redo serially so the carried position follows the reference
        double C0[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < U.natom; a++) scan(a, C0);
    }
    if (bounds_error) { g_ucell_err = "grid box reaches below the cell range (the reference raises BoundsError)"; return false; }
    U.gridn.assign(U.natom, 0);
    U.gl_off.assign(U.natom + 1, 0);
    for (int a = 0; a < U.natom; a++) {
        U.gridn[a] = (int)lists[a].size();
        U.gl_off[a + 1] = U.gl_off[a] + U.gridn[a];
    }
    U.glist.resize(U.gl_off[U.natom]);
    U.clist.resize(U.gl_off[U.natom]);
    for (int a = 0; a < U.natom; a++)
        for (size_t t = 0; t < lists[a].size(); t++) {
            U.glist[U.gl_off[a] + t] = lists[a][t].g;
            U.clist[U.gl_off[a] + t] = lists[a][t].c;
        }
    return true;
}

bool split(UCell &U) {
    const int64_t nloop = U.nb_off[U.natom];
    U.nloop = nloop;
    const int P = U.nprocs;
    if (nloop < P) { g_ucell_err = "number of work items must be at least the number of MPI ranks"; return false; }
    std::vector<double> cum(nloop);
    std::vector<int> oa(nloop), oh(nloop);
    int64_t k = 0;
    double acc = 0.0;
    for (int a = 0; a < U.natom; a++)
        for (int h = 0; h <= U.fnan[a]; h++) {
            const int b = U.natn[U.nb_off[a] + h];
            double w = (double)U.no[a] * (double)U.no[b] * (double)std::min(U.gridn[a], U.gridn[b]);
            if (w < 1.0) w = 1.0;
            acc += w;
            cum[k] = acc; oa[k] = a; oh[k] = h; k++;
        }
    const double total = cum[nloop - 1];
    std::vector<int64_t> first(P + 1);
    int64_t fi = 1;
    for (int r = 1; r <= P - 1; r++) {
        const double target = total * r / P;
        int64_t boundary = std::lower_bound(cum.begin(), cum.end(), target) - cum.begin() + 1;
        const int64_t last_allowed = nloop - (P - r);
        boundary = std::max(fi, std::min(boundary, last_allowed));
        if (boundary < last_allowed) {
            const double prev_err = boundary == fi ? INFINITY : std::fabs(cum[boundary - 2] - target);
            const double cur_err = std::fabs(cum[boundary - 1] - target);
            boundary -= prev_err < cur_err ? 1 : 0;
            boundary = std::max(boundary, fi);
        }
        first[r - 1] = fi;
        fi = boundary + 1;
    }
    first[P - 1] = fi;
    first[P] = nloop + 1;
    std::vector<int64_t> hsize(P, 0);
    for (int r = 0; r < P; r++)
        for (int64_t q = first[r] - 1; q < first[r + 1] - 1; q++)
            hsize[r] += (int64_t)U.no[oa[q]] * U.no[U.natn[U.nb_off[oa[q]] + oh[q]]];
    U.total_hsize = 0;
    for (int r = 0; r < P; r++) {
        if (r == U.rank) U.my_hoff = U.total_hsize;
        U.total_hsize += hsize[r];
    }
    U.my_hsize = hsize[U.rank];
    const int64_t beg = first[U.rank] - 1, end = first[U.rank + 1] - 1;
    U.npair = end - beg;
    U.p_atom.resize(U.npair); U.p_slot.resize(U.npair); U.p_natn.resize(U.npair); U.p_ncn.resize(U.npair);
    U.p_hoff.resize(U.npair + 1);
    int64_t off = 0;
    for (int64_t q = 0; q < U.npair; q++) {
        const int a = oa[beg + q], h = oh[beg + q];
        U.p_atom[q] = a; U.p_slot[q] = h;
        U.p_natn[q] = U.natn[U.nb_off[a] + h];
        U.p_ncn[q] = U.ncn[U.nb_off[a] + h];
        U.p_hoff[q] = off;
        off += (int64_t)U.no[a] * U.no[U.p_natn[q]];
    }
    U.p_hoff[U.npair] = off;
    return true;
}

bool overlaps(UCell &U, int nthreads) {
    const int64_t NN = (int64_t)U.ngrid[0] * U.ngrid[1] * U.ngrid[2];
    const int C = U.cpy, W = 2 * C + 1;
    // Entries covering each home grid point, atoms ascending, list order within an atom.
    std::vector<int64_t> cnt(NN + 1, 0);
    for (int64_t e = 0; e < U.gl_off[U.natom]; e++) cnt[U.glist[e] + 1]++;
    for (int64_t g = 0; g < NN; g++) cnt[g + 1] += cnt[g];
    std::vector<int> ent_atom(U.gl_off[U.natom]), ent_t(U.gl_off[U.natom]);
    {
        std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
        for (int a = 0; a < U.natom; a++)
            for (int t = 0; t < U.gridn[a]; t++) {
                const int64_t pos = fill[U.glist[U.gl_off[a] + t]]++;
                ent_atom[pos] = a;
                ent_t[pos] = t;
            }
    }
    // (atom A, atom B, cell offset) -> rank-local pair index
    std::vector<int> pair_of((size_t)U.natom * U.natom * W * W * W, -1);
    auto key = [&](int a, int b, int i, int j, int k) {
        return (((size_t)a * U.natom + b) * W + (i + C)) * W * W + (size_t)(j + C) * W + (k + C);
    };
    for (int64_t q = 0; q < U.npair; q++) {
        const int *ijk = &U.atv_ijk[3 * (size_t)U.p_ncn[q]];
        pair_of[key(U.p_atom[q], U.p_natn[q], ijk[0], ijk[1], ijk[2])] = (int)q;
    }
    // Count per (thread chunk, pair), then fill in grid order.
    const int nchunk = std::max(1, std::min<int>(nthreads * 4, (int)std::max<int64_t>(1, NN / 4096)));
    std::vector<std::vector<int64_t>> ccount(nchunk, std::vector<int64_t>(U.npair + 1, 0));
    auto visit = [&](int64_t ch, bool fill_mode, std::vector<int64_t> *pos) {
        const int64_t g0 = NN * ch / nchunk, g1 = NN * (ch + 1) / nchunk;
        auto &cc = ccount[ch];
        for (int64_t g = g0; g < g1; g++)
            for (int64_t w = cnt[g]; w < cnt[g + 1]; w++) {      // B entry (list order)
                const int b = ent_atom[w];
                const int cb = U.clist[U.gl_off[b] + ent_t[w]];
                for (int64_t u = cnt[g]; u < cnt[g + 1]; u++) {  // A entry
                    const int a = ent_atom[u];
                    const int ca = U.clist[U.gl_off[a] + ent_t[u]];
                    const int i = U.atv_ijk[3 * (size_t)ca] - U.atv_ijk[3 * (size_t)cb];
                    const int j = U.atv_ijk[3 * (size_t)ca + 1] - U.atv_ijk[3 * (size_t)cb + 1];
                    const int k = U.atv_ijk[3 * (size_t)ca + 2] - U.atv_ijk[3 * (size_t)cb + 2];
                    if (std::abs(i) > C || std::abs(j) > C || std::abs(k) > C) continue;
                    const int q = pair_of[key(a, b, i, j, k)];
                    if (q < 0) continue;
                    if (!fill_mode) cc[q]++;
                    else {
                        const int64_t p = (*pos)[q]++;
                        U.olg1[p] = ent_t[u];
                        U.olg2[p] = ent_t[w];
                    }
                }
            }
    };
    parallel_for(nchunk, nthreads, [&](int64_t ch, int) { visit(ch, false, nullptr); });
    U.numolg.assign(U.npair, 0);
    U.olg_off.assign(U.npair + 1, 0);
    for (int64_t q = 0; q < U.npair; q++) {
        int64_t s = 0;
        for (int ch = 0; ch < nchunk; ch++) s += ccount[ch][q];
        if (s > INT32_MAX) { g_ucell_err = "overlap list longer than 2^31"; return false; }
        U.numolg[q] = (int)s;
        U.olg_off[q + 1] = U.olg_off[q] + s;
    }
    U.olg1.resize(U.olg_off[U.npair]);
    U.olg2.resize(U.olg_off[U.npair]);
    std::vector<std::vector<int64_t>> start(nchunk, std::vector<int64_t>(U.npair));
    for (int64_t q = 0; q < U.npair; q++) {
        int64_t p = U.olg_off[q];
        for (int ch = 0; ch < nchunk; ch++) { start[ch][q] = p; p += ccount[ch][q]; }
    }
    parallel_for(nchunk, nthreads, [&](int64_t ch, int) { visit(ch, true, &start[ch]); });
    return true;
}

}  // namespace

extern "C" {

int lcpao_ucell_build(lcpao_ucell **out, const double *latvecs, int32_t natom, const double *gxyz,
                      const double *cut, const int32_t *no, const int32_t *ngrid, const double *origin,
                      int32_t nprocs, int32_t rank, int32_t nthreads) {
    if (!out || !latvecs || natom <= 0 || !gxyz || !cut || !ngrid || nprocs <= 0 || rank < 0 || rank >= nprocs) {
        g_ucell_err = "invalid argument";
        return LCPAO_EINVAL;
    }
    for (int d = 0; d < 3; d++)
        if (ngrid[d] <= 0) { g_ucell_err = "grid sizes must be positive"; return LCPAO_EINVAL; }
    if ((int64_t)ngrid[0] * ngrid[1] * ngrid[2] > INT32_MAX) { g_ucell_err = "grid larger than 2^31 points"; return LCPAO_ERANGE; }
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    auto *U = new UCell();
    U->natom = natom;
    std::memcpy(U->lat, latvecs, sizeof U->lat);
    std::memcpy(U->ngrid, ngrid, sizeof U->ngrid);
    if (origin) std::memcpy(U->origin, origin, sizeof U->origin);
    else U->origin[0] = U->origin[1] = U->origin[2] = 0.0;
    U->gxyz.assign(gxyz, gxyz + 3 * (size_t)natom);
    U->cut.assign(cut, cut + natom);
    U->no.assign(natom, 0);
    if (no) for (int a = 0; a < natom; a++) U->no[a] = no[a];
    U->nprocs = nprocs;
    U->rank = rank;
    try {
        if (!neighbours(*U, nthreads)) { delete U; return LCPAO_EINVAL; }
        gen_cells(*U);
        U->gridvol = std::fabs(det3(latvecs)) / ((double)ngrid[0] * ngrid[1] * ngrid[2]);
        if (!atoms_grid(*U, nthreads)) { delete U; return LCPAO_EINVAL; }
        if (!split(*U)) { delete U; return LCPAO_EINVAL; }
        if (!overlaps(*U, nthreads)) { delete U; return LCPAO_ERANGE; }
    } catch (const std::bad_alloc &) {
        g_ucell_err = "out of host memory";
        delete U;
        return LCPAO_ERANGE;
    }
    *out = reinterpret_cast<lcpao_ucell *>(U);
    return LCPAO_OK;
}

int lcpao_ucell_scalars(const lcpao_ucell *u, int32_t *cpycell, int64_t *nloop, int64_t *npairs,
                        int64_t *total_hsize, int64_t *my_hsize, int64_t *my_hoff, double *gridvol,
                        int64_t *total_gridn, int64_t *total_olg) {
    if (!u) { g_ucell_err = "null ucell"; return LCPAO_EINVAL; }
    const UCell *U = reinterpret_cast<const UCell *>(u);
    if (cpycell) *cpycell = U->cpy;
    if (nloop) *nloop = U->nloop;
    if (npairs) *npairs = U->npair;
    if (total_hsize) *total_hsize = U->total_hsize;
    if (my_hsize) *my_hsize = U->my_hsize;
    if (my_hoff) *my_hoff = U->my_hoff;
    if (gridvol) *gridvol = U->gridvol;
    if (total_gridn) *total_gridn = U->gl_off[U->natom];
    if (total_olg) *total_olg = U->olg_off[U->npair];
    return LCPAO_OK;
}

int lcpao_ucell_copy(const lcpao_ucell *u, int32_t *fnan, int64_t *nb_off, int32_t *natn, int32_t *ncn,
                     int32_t *gridn, int64_t *gl_off, int32_t *glist, int32_t *clist, int32_t *pair_atom,
                     int32_t *pair_natn, int32_t *pair_ncn, int64_t *pair_hoff, int32_t *numolg,
                     int64_t *olg_off, int32_t *olg1, int32_t *olg2) {
    if (!u) { g_ucell_err = "null ucell"; return LCPAO_EINVAL; }
    const UCell *U = reinterpret_cast<const UCell *>(u);
    auto cp = [](auto *dst, const auto &src) { if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(src[0])); };
    cp(fnan, U->fnan); cp(nb_off, U->nb_off); cp(natn, U->natn); cp(ncn, U->ncn);
    cp(gridn, U->gridn); cp(gl_off, U->gl_off); cp(glist, U->glist); cp(clist, U->clist);
    cp(pair_atom, U->p_atom); cp(pair_natn, U->p_natn); cp(pair_ncn, U->p_ncn); cp(pair_hoff, U->p_hoff);
    cp(numolg, U->numolg); cp(olg_off, U->olg_off); cp(olg1, U->olg1); cp(olg2, U->olg2);
    return LCPAO_OK;
}

const char *lcpao_ucell_last_error(void) { return g_ucell_err.c_str(); }

void lcpao_ucell_destroy(lcpao_ucell *u) { delete reinterpret_cast<UCell *>(u); }

}  // extern "C"
