/*
 * lcpao_oracle.c -- CPU oracle for the LCPAO real-space grid path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference algorithm (LCPAODFT.jl, Julia) used to check the HIP library.
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load it.  The product path never calls it.
 *
 * Every function follows a reference routine operation by operation and
 * cites the file:line it restates.  Paths are relative to the reference
 * repository root.  Build with -ffp-contract=off: Julia never contracts
 * a*b+c into an FMA, and x^2 / x^3 are repeated multiplies (literal_pow).
 *
 * Index conventions used here (all 0-based):
 *   atoms a = 0..Natom-1; neighbour slot h = 0..FNAN[a] (slot 0 = self);
 *   cells c = 0..(2C+1)^3-1 in Generation_ATV order (cell 0 = home);
 *   grid index g = n1*N2*N3 + n2*N3 + n3 (n3 fastest);
 *   table column t = position in the atom's sorted grid list.
 *
 * Pinning (what checks each function):
 *   orc_pao_read            : basis metadata (tests), table golden vectors
 *   orc_ucell_build         : test_UCell.jl:19-50 golden integers
 *   orc_orbitals            : test_Set_Orbs_Grid.jl:22-50 golden values,
 *                             mesh-node known answer, Ylm closed forms
 *   orc_dorbitals           : central differences of orc_orbitals
 *   orc_overlap             : set definition (orc_overlap_bruteforce)
 *   orc_hamiltonian/density : brute-force point-major sums, duality,
 *                             Hermiticity, linearity (tests)
 *   orc_force3              : central differences of the grid energy
 *   orc_hs_matrix           : Hermiticity, k-periodicity, Gamma sum rule,
 *                             duality with orc_dm_collinear (tests)
 *   orc_dm_collinear        : duality with orc_hs_matrix, Gamma Gram blocks
 *   orc_hs_matrix_nc        : linearity reduction of each spin block to
 *                             orc_hs_matrix (real part + im * imaginary part
 *                             at its block position), zero down-up block
 *   orc_dm_nc               : spin-diagonal channels = orc_dm_collinear of
 *                             each spin half; up-down channels and iDM from
 *                             states with C_down = C_up and C_down = i C_up,
 *                             which turn them into orc_dm_collinear outputs
 *                             (tests)
 */
#define _GNU_SOURCE
#include <ctype.h>
#include <dlfcn.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define ORC_PI 3.141592653589793

static char orc_err[512];
const char *orc_last_error(void) { return orc_err; }
static int fail(const char *msg) {
    snprintf(orc_err, sizeof orc_err, "%s", msg);
    return -1;
}

/* ------------------------------------------------------------------ */
/* PAO radial data: Read_PAO.jl:49-76 (basis string), :114-227 (file) */
/* ------------------------------------------------------------------ */
typedef struct {
    int maxl;          /* Spe_MaxL_Basis */
    int num_basis[4];  /* Spe_Num_Basis[l] */
    int total_orbs;    /* Spe_Total_NumOrbs */
    int nmesh;         /* Spe_Num_Mesh_PAO */
    double cut;        /* Spe_Atom_Cut1 */
    int pao_lmax, pao_mul;
    double *xv, *rv;   /* nmesh */
    double *rwf;       /* [(L*pao_mul + mul)*nmesh + i], L<=pao_lmax */
} orc_pao;

/* strip() + whitespace split: returns pointer to k-th (0-based) token */
static const char *nth_token(const char *s, int k, char *buf, size_t n) {
    int idx = -1;
    while (*s) {
        while (*s && isspace((unsigned char)*s)) s++;
        if (!*s) break;
        const char *b = s;
        while (*s && !isspace((unsigned char)*s)) s++;
        if (++idx == k) {
            size_t len = (size_t)(s - b);
            if (len >= n) len = n - 1;
            memcpy(buf, b, len);
            buf[len] = 0;
            return buf;
        }
    }
    return NULL;
}

/* get_ialpha_index, Read_PAO.jl:49-76 */
static int parse_basis(const char *basis, orc_pao *p) {
    size_t n = strlen(basis);
    if (n % 2 == 1 || n < 2 || n > 8) return fail("basis string must be like s2p2d1");
    p->maxl = (int)(n / 2) - 1;
    const char letters[4] = {'s', 'p', 'd', 'f'};
    for (int l = 0; l <= p->maxl; l++) {
        if (basis[2 * l] != letters[l]) return fail("basis string letters must be s,p,d,f in order");
        if (!isdigit((unsigned char)basis[2 * l + 1])) return fail("bad multiplicity");
        p->num_basis[l] = basis[2 * l + 1] - '0';
    }
    p->total_orbs = 0;
    for (int l = 0; l <= p->maxl; l++) p->total_orbs += p->num_basis[l] * (2 * l + 1);
    return 0;
}

int orc_pao_read(const char *path, double cutoff, const char *basis, orc_pao **out) {
    orc_pao *p = calloc(1, sizeof *p);
    if (!p) return fail("oom");
    if (parse_basis(basis, p)) { free(p); return -1; }
    FILE *f = fopen(path, "r");
    if (!f) { free(p); return fail("cannot open PAO file"); }
    char *line = NULL;
    size_t cap = 0;
    char tok[128];
    p->cut = -1.0;
    int po[4] = {0, 0, 0, 0};
    /* first pass: scalar keys, the last matching line wins (Read_PAO.jl:130-158) */
    while (getline(&line, &cap, f) > 0) {
        if (strstr(line, "grid.num.output")) { if (nth_token(line, 1, tok, sizeof tok)) p->nmesh = atoi(tok); po[0] = 1; continue; }
        if (strstr(line, "radial.cutoff.pao")) { if (nth_token(line, 1, tok, sizeof tok)) p->cut = strtod(tok, NULL); po[1] = 1; continue; }
        if (strstr(line, "PAO.Lmax")) { if (nth_token(line, 1, tok, sizeof tok)) p->pao_lmax = atoi(tok); po[2] = 1; continue; }
        if (strstr(line, "PAO.Mul")) { if (nth_token(line, 1, tok, sizeof tok)) p->pao_mul = atoi(tok); po[3] = 1; continue; }
    }
    if (fabs(p->cut - cutoff) > 1e-10) { fclose(f); free(line); free(p); return fail("Spe_Atom_Cut1 mismatch"); }
    if (p->nmesh == 0 || p->cut < 0.0 || p->pao_lmax == 0 || p->pao_mul == 0 || !(po[0] && po[1] && po[2] && po[3])) {
        fclose(f); free(line); free(p); return fail("incomplete PAO header");
    }
    if (p->maxl > p->pao_lmax) { fclose(f); free(line); free(p); return fail("basis l exceeds PAO.Lmax"); }
    for (int l = 0; l <= p->maxl; l++)
        if (p->num_basis[l] > p->pao_mul) { fclose(f); free(line); free(p); return fail("multiplicity exceeds PAO.Mul"); }
    p->xv = calloc((size_t)p->nmesh, sizeof(double));
    p->rv = calloc((size_t)p->nmesh, sizeof(double));
    p->rwf = calloc((size_t)(p->pao_lmax + 1) * p->pao_mul * p->nmesh, sizeof(double));
    /* second pass: valence density mesh (Read_PAO.jl:185-205) */
    rewind(f);
    int got_mesh = 0;
    while (getline(&line, &cap, f) > 0) {
        if (strstr(line, "<valence.charge.density")) {
            for (int i = 0; i < p->nmesh; i++) {
                if (getline(&line, &cap, f) <= 0) break;
                p->xv[i] = strtod(nth_token(line, 0, tok, sizeof tok), NULL);
                p->rv[i] = strtod(nth_token(line, 1, tok, sizeof tok), NULL);
            }
            if (getline(&line, &cap, f) > 0 && strstr(line, "valence.charge.density>")) got_mesh = 1;
            else { got_mesh = 0; break; }
        }
    }
    if (!got_mesh) { fclose(f); free(line); return fail("valence.charge.density block"); }
    /* radial functions per L (Read_Spe_PAO_RWF!, Read_PAO.jl:80-111) */
    for (int L = 0; L <= p->pao_lmax; L++) {
        char mb[64], me[64];
        snprintf(mb, sizeof mb, "<pseudo.atomic.orbitals.L=%d", L);
        snprintf(me, sizeof me, "pseudo.atomic.orbitals.L=%d>", L);
        rewind(f);
        int ok = 0;
        while (getline(&line, &cap, f) > 0) {
            if (strstr(line, mb)) {
                for (int i = 0; i < p->nmesh; i++) {
                    if (getline(&line, &cap, f) <= 0) break;
                    for (int m = 0; m < p->pao_mul; m++) {
                        const char *t = nth_token(line, m + 2, tok, sizeof tok);
                        p->rwf[((size_t)L * p->pao_mul + m) * p->nmesh + i] = t ? strtod(t, NULL) : 0.0;
                    }
                }
                if (getline(&line, &cap, f) > 0 && strstr(line, me)) ok = 1;
                else { ok = 0; break; }
            }
        }
        if (!ok) { fclose(f); free(line); return fail("pseudo.atomic.orbitals block"); }
    }
    fclose(f);
    free(line);
    *out = p;
    return 0;
}

void orc_pao_free(orc_pao *p) {
    if (!p) return;
    free(p->xv); free(p->rv); free(p->rwf); free(p);
}
int orc_pao_info(const orc_pao *p, int *maxl, int *num_basis4, int *total_orbs, int *nmesh, double *cut) {
    *maxl = p->maxl;
    for (int l = 0; l < 4; l++) num_basis4[l] = l <= p->maxl ? p->num_basis[l] : 0;
    *total_orbs = p->total_orbs; *nmesh = p->nmesh; *cut = p->cut;
    return 0;
}
const double *orc_pao_rv(const orc_pao *p) { return p->rv; }
const double *orc_pao_rwf(const orc_pao *p, int l, int mul) {
    return p->rwf + ((size_t)l * p->pao_mul + mul) * p->nmesh;
}

/* ------------------------------------------------------------------ */
/* Geometry: UCell.jl / Generation_ATV.jl / Find_CGrid.jl              */
/* ------------------------------------------------------------------ */
typedef struct {
    int natom, cpy, ncell, nr;     /* nr = 2*cpy+4 : ratv edge */
    double lat[9];                 /* lat[3*i+j] = Latvecs[i+1, j+1] (rows = vectors) */
    int ngrid[3];
    double origin[3];
    double gridvol;
    double *gxyz, *cut;
    int *no;                       /* orbitals per atom (may be 0) */
    double *atv;                   /* ncell x 3 */
    int *atv_ijk;                  /* ncell x 3 */
    int *ratv;                     /* nr^3, [(i*nr + j)*nr + k], i = cell_i + cpy (0-based of Julia's +1) */
    int *fnan;                     /* natom */
    int64_t *nb_off;               /* natom+1 */
    int *natn, *ncn;               /* 0-based atoms, cells */
    double *dis;
    int *gridn;                    /* natom */
    int64_t *gl_off;               /* natom+1 */
    int *glist, *clist;            /* GridListAtom, CellListAtom */
    /* rank-local pair set (split_system_grid) */
    int nprocs, rank;
    int64_t nloop, npair;          /* all pairs / local pairs */
    int *p_atom, *p_slot, *p_natn, *p_ncn;
    int64_t *p_hoff;               /* offset of each local pair block inside the rank slice */
    int64_t total_hsize, my_hsize, my_hoff;  /* my_hoff = MPHks[rank] */
    int64_t *mpi_hsize;            /* nprocs */
    /* overlap lists (Calc_AtomOverlap_Grid) */
    int *numolg;
    int64_t *olg_off;              /* npair+1 */
    int *olg1, *olg2;
} orc_geom;

/* Generation_ATV! (Generation_ATV.jl:1-38) with atv_ijk and ratv */
static void gen_atv(orc_geom *G) {
    int N = G->cpy, nr = G->nr;
    int Rn = 1;
    for (int i = -N; i <= N; i++)
        for (int j = -N; j <= N; j++)
            for (int k = -N; k <= N; k++) {
                int slot;
                if (i == 0 && j == 0 && k == 0) slot = 0;
                else slot = Rn++;
                const double di = (double)i, dj = (double)j, dk = (double)k;
                double *a = G->atv + 3 * slot;
                if (slot == 0) { a[0] = a[1] = a[2] = 0.0; }
                else {
                    a[0] = di * G->lat[0] + dj * G->lat[3] + dk * G->lat[6];
                    a[1] = di * G->lat[1] + dj * G->lat[4] + dk * G->lat[7];
                    a[2] = di * G->lat[2] + dj * G->lat[5] + dk * G->lat[8];
                }
                G->atv_ijk[3 * slot] = i; G->atv_ijk[3 * slot + 1] = j; G->atv_ijk[3 * slot + 2] = k;
                G->ratv[((i + N) * nr + (j + N)) * nr + (k + N)] = slot;
            }
}

static double *atv_for(const double *lat, int N, int *ncell_out) {
    int nc = (2 * N + 1) * (2 * N + 1) * (2 * N + 1);
    double *atv = calloc((size_t)nc * 3, sizeof(double));
    int Rn = 1;
    for (int i = -N; i <= N; i++)
        for (int j = -N; j <= N; j++)
            for (int k = -N; k <= N; k++) {
                if (i == 0 && j == 0 && k == 0) continue;
                const double di = (double)i, dj = (double)j, dk = (double)k;
                double *a = atv + 3 * (Rn++);
                a[0] = di * lat[0] + dj * lat[3] + dk * lat[6];
                a[1] = di * lat[1] + dj * lat[4] + dk * lat[7];
                a[2] = di * lat[2] + dj * lat[5] + dk * lat[8];
            }
    *ncell_out = nc;
    return atv;
}

/* Estimate_Trn_System (UCell.jl:573-608) -> sum of FNAN */
static int64_t estimate_fnan_sum(const orc_geom *G, const double *atv, int ncell, int *fnan) {
    int64_t tot = 0;
    for (int a = 0; a < G->natom; a++) {
        const double *g1 = G->gxyz + 3 * a;
        int cnt = 0;
        for (int b = 0; b < G->natom; b++) {
            const double *g2 = G->gxyz + 3 * b;
            double rcut = G->cut[a] + G->cut[b];
            for (int Rn = 0; Rn < ncell; Rn++) {
                if (a == b && Rn == 0) continue;
                double dx = fabs(g1[0] - g2[0] - atv[3 * Rn]);
                double dy = fabs(g1[1] - g2[1] - atv[3 * Rn + 1]);
                double dz = fabs(g1[2] - g2[2] - atv[3 * Rn + 2]);
                if (dx <= rcut && dy <= rcut && dz <= rcut) {
                    double r = sqrt(dx * dx + dy * dy + dz * dz);
                    if (r <= rcut) cnt++;
                }
            }
        }
        if (fnan) fnan[a] = cnt;
        tot += cnt;
    }
    return tot;
}

/* Get_FNAN + Trn_System (UCell.jl:520-556, 612-672) */
static int get_fnan(orc_geom *G) {
    int po = 0, cpy = 0, count = 1;
    const int countmax = 7;
    int64_t tf = 0;
    while (po == 0 && count < countmax) {
        cpy++;
        int nc;
        double *atv = atv_for(G->lat, cpy, &nc);
        int64_t prev = tf;
        tf = estimate_fnan_sum(G, atv, nc, NULL);
        free(atv);
        if (tf == prev) po = 1;
        count++;
    }
    if (count == countmax || po == 0) return fail("Get_FNAN did not converge");
    G->cpy = cpy;
    int nc;
    double *atv = atv_for(G->lat, cpy, &nc);
    G->fnan = calloc((size_t)G->natom, sizeof(int));
    G->nb_off = calloc((size_t)G->natom + 1, sizeof(int64_t));
    estimate_fnan_sum(G, atv, nc, G->fnan);
    for (int a = 0; a < G->natom; a++) G->nb_off[a + 1] = G->nb_off[a] + G->fnan[a] + 1;
    int64_t tot = G->nb_off[G->natom];
    G->natn = calloc((size_t)tot, sizeof(int));
    G->ncn = calloc((size_t)tot, sizeof(int));
    G->dis = calloc((size_t)tot, sizeof(double));
    for (int a = 0; a < G->natom; a++) {
        const double *g1 = G->gxyz + 3 * a;
        int64_t base = G->nb_off[a];
        int f = 0;
        for (int b = 0; b < G->natom; b++) {
            const double *g2 = G->gxyz + 3 * b;
            double rcut = G->cut[a] + G->cut[b];
            for (int Rn = 0; Rn < nc; Rn++) {
                if (a == b && Rn == 0) {
                    G->natn[base] = a; G->ncn[base] = 0; G->dis[base] = 0.0;
                } else {
                    double dx = fabs(g1[0] - g2[0] - atv[3 * Rn]);
                    double dy = fabs(g1[1] - g2[1] - atv[3 * Rn + 1]);
                    double dz = fabs(g1[2] - g2[2] - atv[3 * Rn + 2]);
                    if (dx <= rcut && dy <= rcut && dz <= rcut) {
                        double r = sqrt(dx * dx + dy * dy + dz * dz);
                        if (r <= rcut) {
                            f++;
                            G->natn[base + f] = b; G->ncn[base + f] = Rn; G->dis[base + f] = r;
                        }
                    }
                }
            }
        }
    }
    free(atv);
    return 0;
}

/* 3x3 inverse by LU with partial pivoting (stand-in for Julia inv, whose
 * LAPACK rounding is not reproduced; only the bounding box depends on it). */
static void inv3(const double *A, double *Ainv) {
    double a[3][3], inv[3][3];
    int perm[3] = {0, 1, 2};
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) a[i][j] = A[3 * i + j];
    for (int k = 0; k < 3; k++) {
        int p = k;
        for (int i = k + 1; i < 3; i++) if (fabs(a[i][k]) > fabs(a[p][k])) p = i;
        if (p != k) {
            for (int j = 0; j < 3; j++) { double t = a[k][j]; a[k][j] = a[p][j]; a[p][j] = t; }
            int t = perm[k]; perm[k] = perm[p]; perm[p] = t;
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

static double det3_lu(const double *A) {
    double a[3][3];
    double sgn = 1.0;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) a[i][j] = A[3 * i + j];
    for (int k = 0; k < 3; k++) {
        int p = k;
        for (int i = k + 1; i < 3; i++) if (fabs(a[i][k]) > fabs(a[p][k])) p = i;
        if (p != k) { for (int j = 0; j < 3; j++) { double t = a[k][j]; a[k][j] = a[p][j]; a[p][j] = t; } sgn = -sgn; }
        for (int i = k + 1; i < 3; i++) {
            double l = a[i][k] / a[k][k];
            for (int j = k + 1; j < 3; j++) a[i][j] -= l * a[k][j];
        }
    }
    return sgn * (a[0][0] * a[1][1] * a[2][2]);
}

/* Find_ln (Find_CGrid.jl:31-49) */
static void find_ln(int N, int n0, int *l0, int *n1) {
    if (n0 < 0) {
        int an = -n0;
        *l0 = -(int)((double)(an - 1) / (double)N) - 1;
        int al = *l0 < 0 ? -*l0 : *l0;
        *n1 = N - (an - N * (al - 1));
    } else if (N <= n0) {
        *l0 = (int)((double)n0 / (double)N);
        *n1 = n0 - N * (*l0);
    } else {
        *l0 = 0;
        *n1 = n0;
    }
}

typedef struct { int g, c; int64_t ord; } gc_entry;
static int cmp_gc(const void *x, const void *y) {
    const gc_entry *a = x, *b = y;
    if (a->g != b->g) return a->g < b->g ? -1 : 1;
    return a->ord < b->ord ? -1 : (a->ord > b->ord);  /* stable */
}

/* Calc_AtomsGrid (UCell.jl:297-411) with Find_CGrids! (Find_CGrid.jl:1-29) */
static int atoms_grid(orc_geom *G) {
    const int N1 = G->ngrid[0], N2 = G->ngrid[1], N3 = G->ngrid[2];
    double glat[9];
    for (int j = 0; j < 3; j++) {
        glat[j] = G->lat[j] / N1;
        glat[3 + j] = G->lat[3 + j] / N2;
        glat[6 + j] = G->lat[6 + j] / N3;
    }
    /* gRecvecs = 2*pi*inv(gLatvecs') */
    double glt[9], ginv[9], grec[9];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) glt[3 * i + j] = glat[3 * j + i];
    inv3(glt, ginv);
    for (int i = 0; i < 9; i++) grec[i] = 2 * ORC_PI * ginv[i];
    static const int kidx[3][2] = {{1, 2}, {2, 0}, {0, 1}};
    G->gridn = calloc((size_t)G->natom, sizeof(int));
    G->gl_off = calloc((size_t)G->natom + 1, sizeof(int64_t));
    gc_entry **lists = calloc((size_t)G->natom, sizeof(gc_entry *));
    double Cxyz[3] = {0.0, 0.0, 0.0};   /* persists across points and atoms, like the reference */
    for (int a = 0; a < G->natom; a++) {
        const double Gx = G->gxyz[3 * a], Gy = G->gxyz[3 * a + 1], Gz = G->gxyz[3 * a + 2];
        double rcut = G->cut[a] + 0.5;
        int nmin[3], nmax[3];
        for (int k = 0; k < 3; k++) {
            int i = kidx[k][0], j = kidx[k][1];
            const double *Li = G->lat + 3 * i, *Lj = G->lat + 3 * j;
            double vx = Li[1] * Lj[2] - Li[2] * Lj[1];
            double vy = Li[2] * Lj[0] - Li[0] * Lj[2];
            double vz = Li[0] * Lj[1] - Li[1] * Lj[0];
            double coef = 1.0 / sqrt(vx * vx + vy * vy + vz * vz);
            vx = vx * coef; vy = vy * coef; vz = vz * coef;
            double Cx = Gx + rcut * vx - G->origin[0];
            double Cy = Gy + rcut * vy - G->origin[1];
            double Cz = Gz + rcut * vz - G->origin[2];
            nmax[k] = (int)((Cx * grec[3 * k] + Cy * grec[3 * k + 1] + Cz * grec[3 * k + 2]) * 0.5 / ORC_PI);
            Cx = Gx - rcut * vx - G->origin[0];
            Cy = Gy - rcut * vy - G->origin[1];
            Cz = Gz - rcut * vz - G->origin[2];
            nmin[k] = (int)((Cx * grec[3 * k] + Cy * grec[3 * k + 1] + Cz * grec[3 * k + 2]) * 0.5 / ORC_PI);
            if (nmax[k] < nmin[k]) { int t = nmin[k]; nmin[k] = nmax[k]; nmax[k] = t; }
        }
        int64_t np = (int64_t)(nmax[0] - nmin[0] + 1) * (nmax[1] - nmin[1] + 1) * (nmax[2] - nmin[2] + 1);
        gc_entry *L = malloc((size_t)(np > 0 ? np : 1) * sizeof(gc_entry));
        int64_t nct = 0;
        rcut = G->cut[a];
        for (int n1 = nmin[0]; n1 <= nmax[0]; n1++)
            for (int n2 = nmin[1]; n2 <= nmax[1]; n2++)
                for (int n3 = nmin[2]; n3 <= nmax[2]; n3++) {
                    int l1, l2, l3, nn1, nn2, nn3;
                    find_ln(N1, n1, &l1, &nn1);
                    find_ln(N2, n2, &l2, &nn2);
                    find_ln(N3, n3, &l3, &nn3);
                    int Rn;
                    /* the reference tests |l+1| against CpyCell (Find_CGrid.jl:136-140) */
                    if (G->cpy < abs(l1 + 1) || G->cpy < abs(l2 + 1) || G->cpy < abs(l3 + 1)) {
                        Rn = 0;  /* Cxyz left stale, as in the reference */
                    } else {
                        int i1 = G->cpy + l1, i2 = G->cpy + l2, i3 = G->cpy + l3;
                        if (i1 < 0 || i2 < 0 || i3 < 0) { free(L); for (int q = 0; q < a; q++) free(lists[q]); free(lists); return fail("ratv index below range (reference would raise BoundsError)"); }
                        Rn = G->ratv[(i1 * G->nr + i2) * G->nr + i3];
                        const double *av = G->atv + 3 * Rn;
                        Cxyz[0] = av[0] + nn1 * glat[0] + nn2 * glat[3] + nn3 * glat[6] + G->origin[0];
                        Cxyz[1] = av[1] + nn1 * glat[1] + nn2 * glat[4] + nn3 * glat[7] + G->origin[1];
                        Cxyz[2] = av[2] + nn1 * glat[2] + nn2 * glat[5] + nn3 * glat[8] + G->origin[2];
                    }
                    int Ng = nn1 * N2 * N3 + nn2 * N3 + nn3;
                    double dx = Cxyz[0] - Gx, dy = Cxyz[1] - Gy, dz = Cxyz[2] - Gz;
                    double R = sqrt(dx * dx + dy * dy + dz * dz);
                    if (R <= rcut) {
                        L[nct].g = Ng; L[nct].c = Rn; L[nct].ord = nct;
                        nct++;
                    }
                }
        qsort(L, (size_t)nct, sizeof(gc_entry), cmp_gc);  /* stable sortperm (UCell.jl:403-407) */
        lists[a] = L;
        G->gridn[a] = (int)nct;
        G->gl_off[a + 1] = G->gl_off[a] + nct;
    }
    int64_t tot = G->gl_off[G->natom];
    G->glist = malloc((size_t)(tot > 0 ? tot : 1) * sizeof(int));
    G->clist = malloc((size_t)(tot > 0 ? tot : 1) * sizeof(int));
    for (int a = 0; a < G->natom; a++) {
        for (int64_t t = 0; t < G->gridn[a]; t++) {
            G->glist[G->gl_off[a] + t] = lists[a][t].g;
            G->clist[G->gl_off[a] + t] = lists[a][t].c;
        }
        free(lists[a]);
    }
    free(lists);
    return 0;
}

/* _split_weighted_contiguous + split_system_grid (UCell.jl:50-132) */
static int split_pairs(orc_geom *G) {
    int64_t nloop = G->nb_off[G->natom];
    G->nloop = nloop;
    int P = G->nprocs;
    if (nloop < P) return fail("number of work items must be at least the number of MPI ranks");
    double *cum = malloc((size_t)nloop * sizeof(double));
    int *oa = malloc((size_t)nloop * sizeof(int)), *oh = malloc((size_t)nloop * sizeof(int));
    int64_t k = 0;
    double acc = 0.0;
    for (int a = 0; a < G->natom; a++)
        for (int h = 0; h <= G->fnan[a]; h++) {
            int b = G->natn[G->nb_off[a] + h];
            double ow = (double)G->no[a] * (double)G->no[b];
            double gw = (double)(G->gridn[a] < G->gridn[b] ? G->gridn[a] : G->gridn[b]);
            double w = ow * gw;
            if (w < 1.0) w = 1.0;
            acc += w;
            cum[k] = acc; oa[k] = a; oh[k] = h; k++;
        }
    double total = cum[nloop - 1];
    int64_t *first = malloc((size_t)(P + 1) * sizeof(int64_t));
    int64_t fi = 1;  /* 1-based like the reference */
    for (int r = 1; r <= P - 1; r++) {
        double target = total * r / P;
        /* searchsortedfirst: first index with cum >= target (1-based) */
        int64_t lo = 0, hi = nloop;
        while (lo < hi) { int64_t mid = (lo + hi) / 2; if (cum[mid] < target) lo = mid + 1; else hi = mid; }
        int64_t boundary = lo + 1;
        int64_t last_allowed = nloop - (P - r);
        if (boundary < fi) boundary = fi;
        if (boundary > last_allowed) boundary = last_allowed;
        if (boundary < last_allowed) {
            double prev_err = boundary == fi ? INFINITY : fabs(cum[boundary - 2] - target);
            double cur_err = fabs(cum[boundary - 1] - target);
            boundary -= prev_err < cur_err ? 1 : 0;
            if (boundary < fi) boundary = fi;
        }
        first[r - 1] = fi;
        fi = boundary + 1;
    }
    first[P - 1] = fi;
    first[P] = nloop + 1;
    int64_t beg = first[G->rank] - 1, end = first[G->rank + 1] - 1;  /* 0-based [beg,end) */
    G->npair = end - beg;
    G->p_atom = malloc((size_t)(G->npair + 1) * sizeof(int));
    G->p_slot = malloc((size_t)(G->npair + 1) * sizeof(int));
    G->p_natn = malloc((size_t)(G->npair + 1) * sizeof(int));
    G->p_ncn = malloc((size_t)(G->npair + 1) * sizeof(int));
    G->p_hoff = malloc((size_t)(G->npair + 1) * sizeof(int64_t));
    G->mpi_hsize = calloc((size_t)P, sizeof(int64_t));
    for (int r = 0; r < P; r++)
        for (int64_t q = first[r] - 1; q < first[r + 1] - 1; q++) {
            int a = oa[q], b = G->natn[G->nb_off[a] + oh[q]];
            G->mpi_hsize[r] += (int64_t)G->no[a] * G->no[b];
        }
    G->total_hsize = 0;
    G->my_hoff = 0;
    for (int r = 0; r < P; r++) {
        if (r == G->rank) G->my_hoff = G->total_hsize;
        G->total_hsize += G->mpi_hsize[r];
    }
    G->my_hsize = G->mpi_hsize[G->rank];
    int64_t off = 0;
    for (int64_t q = 0; q < G->npair; q++) {
        int a = oa[beg + q], h = oh[beg + q];
        G->p_atom[q] = a; G->p_slot[q] = h;
        G->p_natn[q] = G->natn[G->nb_off[a] + h];
        G->p_ncn[q] = G->ncn[G->nb_off[a] + h];
        G->p_hoff[q] = off;
        off += (int64_t)G->no[a] * G->no[G->p_natn[q]];
    }
    G->p_hoff[G->npair] = off;
    free(cum); free(oa); free(oh); free(first);
    return 0;
}

/* Calc_AtomOverlap_Grid (UCell.jl:425-516), the reference's merge restated */
static int overlap_lists(orc_geom *G) {
    int C = G->cpy;
    int maxg = 0;
    for (int a = 0; a < G->natom; a++) if (G->gridn[a] > maxg) maxg = G->gridn[a];
    int *T1 = malloc((size_t)(maxg + 1) * sizeof(int)), *T2 = malloc((size_t)(maxg + 1) * sizeof(int));
    G->numolg = calloc((size_t)G->npair + 1, sizeof(int));
    G->olg_off = calloc((size_t)G->npair + 1, sizeof(int64_t));
    int64_t cap = 1024, used = 0;
    G->olg1 = malloc((size_t)cap * sizeof(int));
    G->olg2 = malloc((size_t)cap * sizeof(int));
    for (int64_t q = 0; q < G->npair; q++) {
        int a = G->p_atom[q], b = G->p_natn[q], cell = G->p_ncn[q];
        const int *GA = G->glist + G->gl_off[a], *CA = G->clist + G->gl_off[a];
        const int *GB = G->glist + G->gl_off[b], *CB = G->clist + G->gl_off[b];
        int l1 = G->atv_ijk[3 * cell], l2 = G->atv_ijk[3 * cell + 1], l3 = G->atv_ijk[3 * cell + 2];
        int nog = -1, Nc = 0;
        int nga = G->gridn[a];
        for (int Nh = 1; Nh <= G->gridn[b]; Nh++) {
            int GNh = GB[Nh - 1], GRh = CB[Nh - 1];
            int lll1 = l1 + G->atv_ijk[3 * GRh], lll2 = l2 + G->atv_ijk[3 * GRh + 1], lll3 = l3 + G->atv_ijk[3 * GRh + 2];
            if (nga > 0 && GA[0] <= GNh) {
                if (GNh == 0) Nc = 0;
                else {
                    while (GNh <= GA[Nc] && Nc != 0) { Nc -= 10; if (Nc < 0) Nc = 0; }
                }
                if (abs(lll1) <= C && abs(lll2) <= C && abs(lll3) <= C) {
                    int GRh1 = G->ratv[((lll1 + C) * G->nr + (lll2 + C)) * G->nr + (lll3 + C)];
                    int po = 0;
                    while (po == 0 && Nc < nga) {
                        int GNc = GA[Nc], GRc = CA[Nc];
                        if (GNc == GNh && GRc == GRh1) {
                            nog++;
                            T1[nog] = Nc; T2[nog] = Nh - 1;
                            po = 1;
                        } else if (GNh < GNc) {
                            po = 1;
                        }
                        Nc++;
                    }
                    Nc--;
                    if (Nc < 0) Nc = 0;
                }
            }
        }
        int n = nog + 1;
        G->numolg[q] = n;
        if (used + n > cap) {
            while (used + n > cap) cap *= 2;
            G->olg1 = realloc(G->olg1, (size_t)cap * sizeof(int));
            G->olg2 = realloc(G->olg2, (size_t)cap * sizeof(int));
        }
        memcpy(G->olg1 + used, T1, (size_t)n * sizeof(int));
        memcpy(G->olg2 + used, T2, (size_t)n * sizeof(int));
        G->olg_off[q] = used;
        used += n;
    }
    G->olg_off[G->npair] = used;
    free(T1); free(T2);
    return 0;
}

void orc_geom_free(orc_geom *G) {
    if (!G) return;
    free(G->gxyz); free(G->cut); free(G->no); free(G->atv); free(G->atv_ijk); free(G->ratv);
    free(G->fnan); free(G->nb_off); free(G->natn); free(G->ncn); free(G->dis);
    free(G->gridn); free(G->gl_off); free(G->glist); free(G->clist);
    free(G->p_atom); free(G->p_slot); free(G->p_natn); free(G->p_ncn); free(G->p_hoff); free(G->mpi_hsize);
    free(G->numolg); free(G->olg_off); free(G->olg1); free(G->olg2);
    free(G);
}

/* UCell(Nspin, Latvecs, Natom, ..., Grid_Origin; Total_NumOrbs) (UCell.jl:210-294).
 * lat: 3x3 row-major, rows are lattice vectors.  gxyz: natom x 3 Cartesian bohr. */
int orc_ucell_build(const double *lat, int natom, const double *gxyz, const double *cut,
                    const int *no, const int *ngrid, const double *origin,
                    int nprocs, int rank, orc_geom **out) {
    if (natom <= 0 || nprocs <= 0 || rank < 0 || rank >= nprocs) return fail("bad arguments");
    orc_geom *G = calloc(1, sizeof *G);
    G->natom = natom;
    memcpy(G->lat, lat, 9 * sizeof(double));
    memcpy(G->ngrid, ngrid, 3 * sizeof(int));
    memcpy(G->origin, origin, 3 * sizeof(double));
    G->gxyz = malloc((size_t)natom * 3 * sizeof(double));
    G->cut = malloc((size_t)natom * sizeof(double));
    G->no = malloc((size_t)natom * sizeof(int));
    memcpy(G->gxyz, gxyz, (size_t)natom * 3 * sizeof(double));
    memcpy(G->cut, cut, (size_t)natom * sizeof(double));
    for (int a = 0; a < natom; a++) G->no[a] = no ? no[a] : 0;
    G->nprocs = nprocs; G->rank = rank;
    if (get_fnan(G)) { orc_geom_free(G); return -1; }
    G->ncell = (2 * G->cpy + 1) * (2 * G->cpy + 1) * (2 * G->cpy + 1);
    G->nr = 2 * G->cpy + 4;
    G->atv = calloc((size_t)G->ncell * 3, sizeof(double));
    G->atv_ijk = calloc((size_t)G->ncell * 3, sizeof(int));
    G->ratv = calloc((size_t)G->nr * G->nr * G->nr, sizeof(int));
    gen_atv(G);
    G->gridvol = fabs(det3_lu(lat)) / ((double)ngrid[0] * ngrid[1] * ngrid[2]);
    if (atoms_grid(G)) { orc_geom_free(G); return -1; }
    if (split_pairs(G)) { orc_geom_free(G); return -1; }
    if (overlap_lists(G)) { orc_geom_free(G); return -1; }
    *out = G;
    return 0;
}

/* accessors for ctypes */
int orc_geom_scalars(const orc_geom *G, int *cpy, int64_t *nloop, int64_t *npair,
                     int64_t *total_hsize, int64_t *my_hsize, int64_t *my_hoff, double *gridvol,
                     int64_t *nolg_total, int64_t *ngrid_total) {
    *cpy = G->cpy; *nloop = G->nloop; *npair = G->npair;
    *total_hsize = G->total_hsize; *my_hsize = G->my_hsize; *my_hoff = G->my_hoff;
    *gridvol = G->gridvol; *nolg_total = G->olg_off[G->npair];
    *ngrid_total = G->gl_off[G->natom];
    return 0;
}
const int *orc_geom_fnan(const orc_geom *G) { return G->fnan; }
const int64_t *orc_geom_nb_off(const orc_geom *G) { return G->nb_off; }
const int *orc_geom_natn(const orc_geom *G) { return G->natn; }
const int *orc_geom_ncn(const orc_geom *G) { return G->ncn; }
const double *orc_geom_atv(const orc_geom *G) { return G->atv; }
const int *orc_geom_atv_ijk(const orc_geom *G) { return G->atv_ijk; }
const int *orc_geom_gridn(const orc_geom *G) { return G->gridn; }
const int64_t *orc_geom_gl_off(const orc_geom *G) { return G->gl_off; }
const int *orc_geom_glist(const orc_geom *G) { return G->glist; }
const int *orc_geom_clist(const orc_geom *G) { return G->clist; }
const int *orc_geom_pair_atom(const orc_geom *G) { return G->p_atom; }
const int *orc_geom_pair_slot(const orc_geom *G) { return G->p_slot; }
const int *orc_geom_pair_natn(const orc_geom *G) { return G->p_natn; }
const int *orc_geom_pair_ncn(const orc_geom *G) { return G->p_ncn; }
const int64_t *orc_geom_pair_hoff(const orc_geom *G) { return G->p_hoff; }
const int64_t *orc_geom_mpi_hsize(const orc_geom *G) { return G->mpi_hsize; }
const int *orc_geom_numolg(const orc_geom *G) { return G->numolg; }
const int64_t *orc_geom_olg_off(const orc_geom *G) { return G->olg_off; }
const int *orc_geom_olg1(const orc_geom *G) { return G->olg1; }
const int *orc_geom_olg2(const orc_geom *G) { return G->olg2; }

/* Set-definition overlap (independent of the merge): for pair (A,B,cell),
 * every B entry (g, cB) whose partner (g, cB+cell) is in A's list. */
int orc_overlap_bruteforce(const orc_geom *G, int64_t q, int *t1, int *t2, int64_t cap, int64_t *n_out) {
    int a = G->p_atom[q], b = G->p_natn[q], cell = G->p_ncn[q], C = G->cpy;
    const int *GA = G->glist + G->gl_off[a], *CA = G->clist + G->gl_off[a];
    const int *GB = G->glist + G->gl_off[b], *CB = G->clist + G->gl_off[b];
    int64_t n = 0;
    for (int tb = 0; tb < G->gridn[b]; tb++) {
        int l1 = G->atv_ijk[3 * cell] + G->atv_ijk[3 * CB[tb]];
        int l2 = G->atv_ijk[3 * cell + 1] + G->atv_ijk[3 * CB[tb] + 1];
        int l3 = G->atv_ijk[3 * cell + 2] + G->atv_ijk[3 * CB[tb] + 2];
        if (abs(l1) > C || abs(l2) > C || abs(l3) > C) continue;
        int want = G->ratv[((l1 + C) * G->nr + (l2 + C)) * G->nr + (l3 + C)];
        for (int ta = 0; ta < G->gridn[a]; ta++) {
            if (GA[ta] == GB[tb] && CA[ta] == want) {
                if (n < cap) { t1[n] = ta; t2[n] = tb; }
                n++;
                break;
            }
        }
    }
    *n_out = n;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Orbital tables: xyz_to_spherical.jl, Ylm_real.jl,                   */
/* Set_Orbitals_Grid.jl:45-277, Set_dOrbitals_Grid.jl:23-362           */
/* ------------------------------------------------------------------ */
static void xyz_to_spherical(double x, double y, double z, double *r_, double *th_, double *ph_) {
    const double Min_r = 1e-14;
    double dum = x * x + y * y;
    double r = sqrt(dum + z * z);
    double r1 = sqrt(dum);
    double theta, phi;
    if (Min_r <= r) {
        double sz = (z > 0.0) ? 1.0 : ((z < 0.0) ? -1.0 : z);
        double dum1 = (r < fabs(z)) ? sz : z / r;
        theta = acos(dum1);
        if (Min_r <= r1) {
            double sy = (y > 0.0) ? 1.0 : ((y < 0.0) ? -1.0 : y);
            double d2 = (r1 < fabs(y)) ? sy : y / r1;
            if (x >= 0.0) phi = asin(d2);
            else phi = ORC_PI - asin(d2);
        } else {
            phi = 0.0;
        }
    } else {
        theta = 0.5 * ORC_PI;
        phi = 0.0;
    }
    *r_ = r; *th_ = theta; *ph_ = phi;
}

/* Ylm_real (Ylm_real.jl:1-53) */
double orc_ylm_real(int l, int m, double theta, double phi) {
    double siQ = sin(theta), coQ = cos(theta), siP = sin(phi), coP = cos(phi);
    if (l == 0) return 0.5 * sqrt(1.0 / ORC_PI);
    if (l == 1) {
        if (m == -1) return sqrt(3.0 / 4.0 / ORC_PI) * siQ * siP;
        if (m == 0) return sqrt(3.0 / 4.0 / ORC_PI) * coQ;
        return sqrt(3.0 / 4.0 / ORC_PI) * siQ * coP;
    }
    if (l == 2) {
        if (m == -2) return 0.5 * sqrt(15.0 / ORC_PI) * (siQ * siQ) * siP * coP;
        if (m == -1) return 0.5 * sqrt(15.0 / ORC_PI) * siQ * coQ * siP;
        if (m == 0) return 0.25 * sqrt(5.0 / ORC_PI) * (3 * (coQ * coQ) - 1);
        if (m == 1) return 0.5 * sqrt(15.0 / ORC_PI) * siQ * coQ * coP;
        return 0.25 * sqrt(15.0 / ORC_PI) * (siQ * siQ) * (1 - 2 * (siP * siP));
    }
    if (l == 3) {
        if (m == -3) return 0.25 * sqrt(35.0 / 2.0 / ORC_PI) * (siQ * siQ * siQ) * (3 * siP - 4 * (siP * siP * siP));
        if (m == -2) return 0.5 * sqrt(105.0 / ORC_PI) * (siQ * siQ) * coQ * siP * coP;
        if (m == -1) return 0.25 * sqrt(21.0 / 2.0 / ORC_PI) * siQ * (5 * (coQ * coQ) - 1) * siP;
        if (m == 0) return 0.25 * sqrt(7.0 / ORC_PI) * (5 * (coQ * coQ * coQ) - 3 * coQ);
        if (m == 1) return 0.25 * sqrt(21.0 / 2.0 / ORC_PI) * siQ * (5 * (coQ * coQ) - 1) * coP;
        if (m == 2) return 0.25 * sqrt(105.0 / ORC_PI) * (siQ * siQ) * coQ * ((coP * coP) - (siP * siP));
        return 0.25 * sqrt(35.0 / 2.0 / ORC_PI) * (siQ * siQ * siQ) * (4 * (coP * coP * coP) - 3 * coP);
    }
    return NAN;
}

/* dYlmdtheta_real (Ylm_real.jl:56-108) */
double orc_dylm_dtheta(int l, int m, double theta, double phi) {
    double siQ = sin(theta), coQ = cos(theta), siP = sin(phi), coP = cos(phi);
    if (l == 0) return 0.0;
    if (l == 1) {
        if (m == -1) return sqrt(3.0 / 4.0 / ORC_PI) * coQ * siP;
        if (m == 0) return -sqrt(3.0 / 4.0 / ORC_PI) * siQ;
        return sqrt(3.0 / 4.0 / ORC_PI) * coQ * coP;
    }
    if (l == 2) {
        if (m == -2) return sqrt(15.0 / ORC_PI) * siQ * coQ * siP * coP;
        if (m == -1) return 0.5 * sqrt(15.0 / ORC_PI) * (1 - 2 * (siQ * siQ)) * siP;
        if (m == 0) return -1.5 * sqrt(5.0 / ORC_PI) * siQ * coQ;
        if (m == 1) return 0.5 * sqrt(15.0 / ORC_PI) * (1 - 2 * (siQ * siQ)) * coP;
        return 0.5 * sqrt(15.0 / ORC_PI) * siQ * coQ * (1 - 2 * (siP * siP));
    }
    if (l == 3) {
        if (m == -3) return 0.75 * sqrt(35.0 / 2.0 / ORC_PI) * (siQ * siQ) * coQ * siP * (3 - 4 * (siP * siP));
        if (m == -2) return 0.5 * sqrt(105.0 / ORC_PI) * coP * siP * siQ * (2 * (coQ * coQ) - (siQ * siQ));
        if (m == -1) return 0.25 * sqrt(21.0 / 2.0 / ORC_PI) * siP * coQ * (15 * (coQ * coQ) - 11);
        if (m == 0) return 0.25 * sqrt(7.0 / ORC_PI) * siQ * (-15 * (coQ * coQ) + 3);
        if (m == 1) return 0.25 * sqrt(21.0 / 2.0 / ORC_PI) * coP * coQ * (15 * (coQ * coQ) - 11);
        if (m == 2) return 0.25 * sqrt(105.0 / ORC_PI) * ((coP * coP) - (siP * siP)) * siQ * (2 * (coQ * coQ) - (siQ * siQ));
        return 0.75 * sqrt(35.0 / 2.0 / ORC_PI) * coP * coQ * (siQ * siQ) * (-3 + 4 * (coP * coP));
    }
    return NAN;
}

/* dYlmdphi_real (Ylm_real.jl:111-162) */
double orc_dylm_dphi(int l, int m, double theta, double phi) {
    double siQ = sin(theta), coQ = cos(theta), siP = sin(phi), coP = cos(phi);
    if (l == 0) return 0.0;
    if (l == 1) {
        if (m == -1) return sqrt(3.0 / 4.0 / ORC_PI) * coP;
        if (m == 0) return 0.0;
        return -sqrt(3.0 / 4.0 / ORC_PI) * siP;
    }
    if (l == 2) {
        if (m == -2) return 0.5 * sqrt(15.0 / ORC_PI) * siQ * (1 - 2 * (siP * siP));
        if (m == -1) return 0.5 * sqrt(15.0 / ORC_PI) * coQ * coP;
        if (m == 0) return 0.0;
        if (m == 1) return -0.5 * sqrt(15.0 / ORC_PI) * coQ * siP;
        return -sqrt(15.0 / ORC_PI) * siQ * siP * coP;
    }
    if (l == 3) {
        if (m == -3) return 0.75 * sqrt(35.0 / 2.0 / ORC_PI) * coP * (siQ * siQ) * (1 - 4 * (siP * siP));
        if (m == -2) return 0.5 * sqrt(105.0 / ORC_PI) * coQ * siQ * ((coP * coP) - (siP * siP));
        if (m == -1) return 0.25 * sqrt(21.0 / 2.0 / ORC_PI) * coP * (5 * (coQ * coQ) - 1);
        if (m == 0) return 0.0;
        if (m == 1) return 0.25 * sqrt(21.0 / 2.0 / ORC_PI) * siP * (-5 * (coQ * coQ) + 1);
        if (m == 2) return -sqrt(105.0 / ORC_PI) * coP * coQ * siP * siQ;
        return 0.75 * sqrt(35.0 / 2.0 / ORC_PI) * siP * (siQ * siQ) * (1 - 4 * (coP * coP));
    }
    return NAN;
}

/* m order within each l (Set_Orbitals_Grid.jl:243-267) */
static const int M_ORDER[4][7] = {
    {0}, {1, -1, 0}, {0, 2, -2, 1, -1}, {0, 1, -1, 2, -2, 3, -3}};

/* Set_Orbitals_Grid! (Set_Orbitals_Grid.jl:45-277) for every atom.
 * paos[spe], atom2spe[a]; out[a] points to NO_a x GridN_a (column-major,
 * orbital fastest).  The PAO's basis must match geometry's no[a]. */
int orc_orbitals(const orc_geom *G, const orc_pao *const *paos, const int *atom2spe, double *const *out) {
    const int N1 = G->ngrid[0], N2 = G->ngrid[1], N3 = G->ngrid[2];
    double glat[9];
    for (int j = 0; j < 3; j++) { glat[j] = G->lat[j] / N1; glat[3 + j] = G->lat[3 + j] / N2; glat[6 + j] = G->lat[6 + j] / N3; }
    for (int a = 0; a < G->natom; a++) {
        const orc_pao *p = paos[atom2spe[a]];
        if (p->total_orbs != G->no[a]) return fail("orbital count mismatch");
        const double *RV = p->rv;
        const int nm = p->nmesh;
        double RF[4][16], AF[4][7];
        memset(AF, 0, sizeof AF);
        for (int t = 0; t < G->gridn[a]; t++) {
            int GNc = G->glist[G->gl_off[a] + t];
            int GRc = G->clist[G->gl_off[a] + t];
            int n1 = GNc / (N2 * N3);
            int n2 = (GNc - n1 * N2 * N3) / N3;
            int n3 = GNc - n1 * N2 * N3 - n2 * N3;
            double Cx = n1 * glat[0] + n2 * glat[3] + n3 * glat[6] + G->origin[0];
            double Cy = n1 * glat[1] + n2 * glat[4] + n3 * glat[7] + G->origin[1];
            double Cz = n1 * glat[2] + n2 * glat[5] + n3 * glat[8] + G->origin[2];
            double x = Cx + G->atv[3 * GRc] - G->gxyz[3 * a];
            double y = Cy + G->atv[3 * GRc + 1] - G->gxyz[3 * a + 1];
            double z = Cz + G->atv[3 * GRc + 2] - G->gxyz[3 * a + 2];
            double R, theta, phi;
            xyz_to_spherical(x, y, z, &R, &theta, &phi);
            int po = 0;
            for (int l = 0; l <= p->maxl; l++) for (int q = 0; q < p->num_basis[l]; q++) RF[l][q] = 0.0;
            if (RV[nm - 1] < R) {
                po = 1;
            } else if (R < RV[0]) {
                const int m = 5 - 1;  /* Julia m = 5 */
                double rm = RV[m];
                double h1 = RV[m - 1] - RV[m - 2], h2 = RV[m] - RV[m - 1], h3 = RV[m + 1] - RV[m];
                double x1 = rm - RV[m - 1], x2 = rm - RV[m];
                double y1 = x1 / h2, y2 = x2 / h2, y12 = y1 * y1, y22 = y2 * y2;
                double dum = h1 + h2;
                double dum1 = h1 / h2 / dum, dum2 = h2 / h1 / dum;
                dum = h2 + h3;
                double dum3 = h2 / h3 / dum, dum4 = h3 / h2 / dum;
                for (int l = 0; l <= p->maxl; l++)
                    for (int q = 0; q < p->num_basis[l]; q++) {
                        const double *W = orc_pao_rwf(p, l, q);
                        double f1 = W[m - 2], f2 = W[m - 1], f3 = W[m], f4 = W[m + 1];
                        double d = f3 - f2;
                        double g1 = d * dum1 + (f2 - f1) * dum2;
                        double g2 = (f4 - f3) * dum3 + d * dum4;
                        double A = 3 * f2 + h2 * g1, B = 2 * f2 + h2 * g1;
                        double Cc = 3 * f3 - h2 * g2, D = 2 * f3 - h2 * g2;
                        double E = A + B * y2, F = Cc - D * y1;
                        double f = y22 * E + y12 * F;
                        double df = (2 * y2 / h2) * E + y22 * B / h2 + (2 * y1 / h2) * F - y12 * D / h2;
                        double ca, cb, cc, cd;
                        if (l == 0) { ca = 0.0; cb = 0.5 * df / rm; cc = 0.0; cd = f - cb * (rm * rm); }
                        else if (l == 1) { ca = (rm * df - f) / (2 * (rm * rm * rm)); cb = 0.0; cc = df - 3 * ca * (rm * rm); cd = 0.0; }
                        else { cb = (3 * f - rm * df) / (rm * rm); ca = (f - cb * (rm * rm)) / (rm * rm * rm); cc = 0.0; cd = 0.0; }
                        RF[l][q] = ca * (R * R * R) + cb * (R * R) + cc * R + cd;
                    }
            } else {
                int mp_min = 1, mp_max = nm;  /* 1-based */
                while (mp_max - mp_min != 1) {
                    int mm = (mp_min + mp_max) / 2;
                    if (RV[mm - 1] < R) mp_min = mm; else mp_max = mm;
                }
                int m = mp_max;  /* 1-based */
                /* the reference reads RWF[m-2] and RWF[m+1] (1-based) first: out of range raises there */
                if (m < 3 || m + 1 > nm) return fail("radial mesh index out of range (reference raises BoundsError)");
                double h1 = RV[m - 2] - RV[m - 3], h2 = RV[m - 1] - RV[m - 2], h3 = RV[m] - RV[m - 1];
                double x1 = R - RV[m - 2], x2 = R - RV[m - 1];
                double y1 = x1 / h2, y2 = x2 / h2, y12 = y1 * y1, y22 = y2 * y2;
                double dum = h1 + h2;
                double dum1 = h1 / h2 / dum, dum2 = h2 / h1 / dum;
                dum = h2 + h3;
                double dum3 = h2 / h3 / dum, dum4 = h3 / h2 / dum;
                for (int l = 0; l <= p->maxl; l++)
                    for (int q = 0; q < p->num_basis[l]; q++) {
                        const double *W = orc_pao_rwf(p, l, q);
                        double f1 = W[m - 3], f2 = W[m - 2], f3 = W[m - 1], f4 = W[m];
                        if (m == 1) { h1 = -(h2 + h3); f1 = f4; }
                        else if (m == nm - 1) { h3 = -(h1 + h2); f4 = f1; }
                        double d = f3 - f2;
                        double g1 = d * dum1 + (f2 - f1) * dum2;
                        double g2 = (f4 - f3) * dum3 + d * dum4;
                        double f = (y22 * (3 * f2 + h2 * g1 + (2 * f2 + h2 * g1) * y2) + y12 * (3 * f3 - h2 * g2 - (2 * f3 - h2 * g2) * y1));
                        RF[l][q] = f;
                    }
            }
            if (po == 0)
                for (int l = 0; l <= p->maxl; l++)
                    for (int mi = 0; mi < 2 * l + 1; mi++) AF[l][mi] = orc_ylm_real(l, M_ORDER[l][mi], theta, phi);
            double *col = out[a] + (size_t)t * G->no[a];
            int ist = 0;
            for (int l = 0; l <= p->maxl; l++)
                for (int q = 0; q < p->num_basis[l]; q++)
                    for (int mi = 0; mi < 2 * l + 1; mi++) col[ist++] = RF[l][q] * AF[l][mi];
        }
    }
    return 0;
}

/* Set_dOrbitals_Grid! (Set_dOrbitals_Grid.jl:23-362).  out{x,y,z}[a]: NO_a x GridN_a. */
int orc_dorbitals(const orc_geom *G, const orc_pao *const *paos, const int *atom2spe,
                  double *const *outx, double *const *outy, double *const *outz) {
    const int N1 = G->ngrid[0], N2 = G->ngrid[1], N3 = G->ngrid[2];
    double glat[9];
    for (int j = 0; j < 3; j++) { glat[j] = G->lat[j] / N1; glat[3 + j] = G->lat[3 + j] / N2; glat[6 + j] = G->lat[6 + j] / N3; }
    const double Rmin = 1.0e-13;
    double dRx = 0.0, dRy = 0.0, dRz = 0.0, dQx = 0.0, dQy = 0.0, dQz = 0.0, dPx = 0.0, dPy = 0.0, dPz = 0.0;
    double RF[4][16], dRF[4][16], AF[4][7], dAFQ[4][7], dAFP[4][7];
    memset(AF, 0, sizeof AF); memset(dAFQ, 0, sizeof dAFQ); memset(dAFP, 0, sizeof dAFP);
    memset(RF, 0, sizeof RF); memset(dRF, 0, sizeof dRF);
    for (int a = 0; a < G->natom; a++) {
        const orc_pao *p = paos[atom2spe[a]];
        if (p->total_orbs != G->no[a]) return fail("orbital count mismatch");
        const double *RV = p->rv;
        const int nm = p->nmesh;
        for (int t = 0; t < G->gridn[a]; t++) {
            int GNc = G->glist[G->gl_off[a] + t];
            int GRc = G->clist[G->gl_off[a] + t];
            int n1 = GNc / (N2 * N3);
            int n2 = (GNc - n1 * N2 * N3) / N3;
            int n3 = GNc - n1 * N2 * N3 - n2 * N3;
            double Cx = n1 * glat[0] + n2 * glat[3] + n3 * glat[6] + G->origin[0];
            double Cy = n1 * glat[1] + n2 * glat[4] + n3 * glat[7] + G->origin[1];
            double Cz = n1 * glat[2] + n2 * glat[5] + n3 * glat[8] + G->origin[2];
            double x = Cx + G->atv[3 * GRc] - G->gxyz[3 * a];
            double y = Cy + G->atv[3 * GRc + 1] - G->gxyz[3 * a + 1];
            double z = Cz + G->atv[3 * GRc + 2] - G->gxyz[3 * a + 2];
            double R, theta, phi;
            xyz_to_spherical(x, y, z, &R, &theta, &phi);
            if (R < Rmin) {
                x = x + Rmin; y = y + Rmin; z = z + Rmin;
                xyz_to_spherical(x, y, z, &R, &theta, &phi);
            }
            int po = 0;
            if (RV[nm - 1] < R) {
                for (int l = 0; l <= p->maxl; l++) for (int q = 0; q < p->num_basis[l]; q++) { RF[l][q] = 0.0; dRF[l][q] = 0.0; }
                po = 1;
            } else if (R < RV[0]) {
                const int m = 5 - 1;
                double rm = RV[m];
                double h1 = RV[m - 1] - RV[m - 2], h2 = RV[m] - RV[m - 1], h3 = RV[m + 1] - RV[m];
                double x1 = rm - RV[m - 1], x2 = rm - RV[m];
                double y1 = x1 / h2, y2 = x2 / h2, y12 = y1 * y1, y22 = y2 * y2;
                double dum = h1 + h2;
                double dum1 = h1 / h2 / dum, dum2 = h2 / h1 / dum;
                dum = h2 + h3;
                double dum3 = h2 / h3 / dum, dum4 = h3 / h2 / dum;
                for (int l = 0; l <= p->maxl; l++)
                    for (int q = 0; q < p->num_basis[l]; q++) {
                        const double *W = orc_pao_rwf(p, l, q);
                        double f1 = W[m - 2], f2 = W[m - 1], f3 = W[m], f4 = W[m + 1];
                        double d = f3 - f2;
                        double g1 = d * dum1 + (f2 - f1) * dum2;
                        double g2 = (f4 - f3) * dum3 + d * dum4;
                        double A = 3 * f2 + h2 * g1, B = 2 * f2 + h2 * g1;
                        double Cc = 3 * f3 - h2 * g2, D = 2 * f3 - h2 * g2;
                        double E = A + B * y2, F = Cc - D * y1;
                        double f = y22 * E + y12 * F;
                        double df = (2 * y2 * E + y22 * B + 2.0 * y1 * F - y12 * D) / h2;
                        double ca, cb, cc, cd;
                        if (l == 0) { ca = 0.0; cb = 0.5 * df / rm; cc = 0.0; cd = f - cb * (rm * rm); }
                        else if (l == 1) { ca = (rm * df - f) / (2 * (rm * rm * rm)); cb = 0.0; cc = df - 3 * ca * (rm * rm); cd = 0.0; }
                        else { cb = (3 * f - rm * df) / (rm * rm); ca = (f - cb * (rm * rm)) / (rm * rm * rm); cc = 0.0; cd = 0.0; }
                        RF[l][q] = ca * (R * R * R) + cb * (R * R) + cc * R + cd;
                        dRF[l][q] = 3 * ca * (R * R) + 2 * cb * R + cc;
                    }
            } else {
                int mp_min = 1, mp_max = nm;
                while (mp_max - mp_min != 1) {
                    int mm = (mp_min + mp_max) / 2;
                    if (RV[mm - 1] < R) mp_min = mm; else mp_max = mm;
                }
                int m = mp_max;
                if (m < 3 || m + 1 > nm) return fail("radial mesh index out of range (reference raises BoundsError)");
                double h1 = RV[m - 2] - RV[m - 3], h2 = RV[m - 1] - RV[m - 2], h3 = RV[m] - RV[m - 1];
                double x1 = R - RV[m - 2], x2 = R - RV[m - 1];
                double y1 = x1 / h2, y2 = x2 / h2, y12 = y1 * y1, y22 = y2 * y2;
                double dum = h1 + h2;
                double dum1 = h1 / h2 / dum, dum2 = h2 / h1 / dum;
                dum = h2 + h3;
                double dum3 = h2 / h3 / dum, dum4 = h3 / h2 / dum;
                for (int l = 0; l <= p->maxl; l++)
                    for (int q = 0; q < p->num_basis[l]; q++) {
                        const double *W = orc_pao_rwf(p, l, q);
                        double f1 = W[m - 3], f2 = W[m - 2], f3 = W[m - 1], f4 = W[m];
                        if (m == 2) { h1 = -(h2 + h3); f1 = f4; }
                        else if (m == nm) { h3 = -(h1 + h2); f4 = f1; }
                        double d = f3 - f2;
                        double g1 = d * dum1 + (f2 - f1) * dum2;
                        double g2 = (f4 - f3) * dum3 + d * dum4;
                        double A = 3 * f2 + h2 * g1, B = 2 * f2 + h2 * g1;
                        double Cc = 3 * f3 - h2 * g2, D = 2 * f3 - h2 * g2;
                        double E = A + B * y2, F = Cc - D * y1;
                        double f = y22 * E + y12 * F;
                        double df = (2 * y2 * E + y22 * B + 2.0 * y1 * F - y12 * D) / h2;
                        RF[l][q] = f;
                        dRF[l][q] = df;
                    }
            }
            if (po == 0) {
                double siQ = sin(theta), coQ = cos(theta), siP = sin(phi), coP = cos(phi);
                dRx = siQ * coP; dRy = siQ * siP; dRz = coQ;
                if (R > Rmin) {
                    dQx = coQ * coP / R; dQy = coQ * siP / R; dQz = -siQ / R;
                    dPx = -siP / R; dPy = coP / R; dPz = 0.0;
                } else {
                    dQx = dQy = dQz = dPx = dPy = dPz = 0.0;
                }
                for (int l = 0; l <= p->maxl; l++)
                    for (int mi = 0; mi < 2 * l + 1; mi++) {
                        int mm = M_ORDER[l][mi];
                        AF[l][mi] = orc_ylm_real(l, mm, theta, phi);
                        dAFQ[l][mi] = l == 0 ? 0.0 : orc_dylm_dtheta(l, mm, theta, phi);
                        dAFP[l][mi] = l == 0 ? 0.0 : orc_dylm_dphi(l, mm, theta, phi);
                    }
            }
            int NO = G->no[a];
            double *cx = outx[a] + (size_t)t * NO, *cy = outy[a] + (size_t)t * NO, *cz = outz[a] + (size_t)t * NO;
            int ist = 0;
            for (int l = 0; l <= p->maxl; l++)
                for (int q = 0; q < p->num_basis[l]; q++)
                    for (int mi = 0; mi < 2 * l + 1; mi++) {
                        double dchidr = dRF[l][q] * AF[l][mi];
                        double dchidtheta = RF[l][q] * dAFQ[l][mi];
                        double dchidphi = RF[l][q] * dAFP[l][mi];
                        cx[ist] = -dRx * dchidr - dQx * dchidtheta - dPx * dchidphi;
                        cy[ist] = -dRy * dchidr - dQy * dchidtheta - dPy * dchidphi;
                        cz[ist] = -dRz * dchidr - dQz * dchidtheta - dPz * dchidphi;
                        ist++;
                    }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Blocked contractions in reference order                             */
/* ------------------------------------------------------------------ */
#define ORC_BLOCK 512  /* density_block_size = ham_block_size (LCPAODFT.jl:54-55) */

/* _Calc_Hamiltonian_Blocked! (Hamiltonian.jl:384-429) driven by
 * Calc_MatrixElements_dVH_Vxc_{off,on,nc}! (Hamiltonian.jl:235-380).
 * tab[a]: NO_a x GridN_a; V[c]: NN; H[c]: my_hsize (overwritten, x GridVol).
 * With absH != NULL, also accumulates GridVol * sum |phi_i V phi_j| (the
 * magnitude used by the parity tolerance). */
int orc_hamiltonian(const orc_geom *G, const double *const *tab, int nch, const double *const *V,
                    double *const *H, double *const *absH) {
    int maxno = 0;
    for (int a = 0; a < G->natom; a++) if (G->no[a] > maxno) maxno = G->no[a];
    double *ob1 = malloc((size_t)maxno * ORC_BLOCK * sizeof(double));
    double *ob2 = malloc((size_t)maxno * ORC_BLOCK * sizeof(double));
    double *wb = malloc((size_t)maxno * ORC_BLOCK * sizeof(double));
    double *ps = malloc((size_t)maxno * maxno * sizeof(double));
    for (int c = 0; c < nch; c++) {
        memset(H[c], 0, (size_t)G->my_hsize * sizeof(double));
        if (absH) memset(absH[c], 0, (size_t)G->my_hsize * sizeof(double));
    }
    for (int64_t q = 0; q < G->npair; q++) {
        int a = G->p_atom[q], b = G->p_natn[q];
        int NO0 = G->no[a], NO1 = G->no[b];
        const double *o1 = tab[a], *o2 = tab[b];
        const int *L1 = G->olg1 + G->olg_off[q], *L2 = G->olg2 + G->olg_off[q];
        const int *gl = G->glist + G->gl_off[a];
        int n = G->numolg[q];
        int64_t hst = G->p_hoff[q];
        for (int bs = 0; bs < n; bs += ORC_BLOCK) {
            int pts = n - bs < ORC_BLOCK ? n - bs : ORC_BLOCK;
            for (int k = 0; k < pts; k++) {
                int t1 = L1[bs + k], t2 = L2[bs + k];
                for (int i = 0; i < NO0; i++) ob1[(size_t)k * NO0 + i] = o1[(size_t)t1 * NO0 + i];
                for (int j = 0; j < NO1; j++) ob2[(size_t)k * NO1 + j] = o2[(size_t)t2 * NO1 + j];
            }
            for (int c = 0; c < nch; c++) {
                for (int k = 0; k < pts; k++) {
                    double v = V[c][gl[L1[bs + k]]];
                    for (int i = 0; i < NO0; i++) wb[(size_t)k * NO0 + i] = ob1[(size_t)k * NO0 + i] * v;
                }
                /* product[jst, ist] = sum_k ob2[jst,k] * wb[ist,k] */
                for (int i = 0; i < NO0; i++)
                    for (int j = 0; j < NO1; j++) {
                        double s = 0.0;
                        for (int k = 0; k < pts; k++) s += ob2[(size_t)k * NO1 + j] * wb[(size_t)k * NO0 + i];
                        ps[(size_t)i * NO1 + j] = s;
                    }
                for (int i = 0; i < NO0; i++)
                    for (int j = 0; j < NO1; j++) H[c][hst + (int64_t)i * NO1 + j] += ps[(size_t)i * NO1 + j];
                if (absH) {
                    for (int i = 0; i < NO0; i++)
                        for (int j = 0; j < NO1; j++) {
                            double s = 0.0;
                            for (int k = 0; k < pts; k++) s += fabs(ob2[(size_t)k * NO1 + j] * wb[(size_t)k * NO0 + i]);
                            absH[c][hst + (int64_t)i * NO1 + j] += s;
                        }
                }
            }
        }
    }
    for (int c = 0; c < nch; c++) {
        for (int64_t h = 0; h < G->my_hsize; h++) H[c][h] *= G->gridvol;
        if (absH) for (int64_t h = 0; h < G->my_hsize; h++) absH[c][h] *= G->gridvol;
    }
    free(ob1); free(ob2); free(wb); free(ps);
    return 0;
}

/* Set_Density_Grid_{nonpol,pol,nc}! (Set_Density_Grid.jl:60-240) with
 * _Calc_Density_Blocked! (:268-304).  DM[s]: global flat array (the rank's
 * slice starts at my_hoff).  rho[s]: NN (overwritten, rank-local partial).
 * With absrho != NULL, also accumulates sum |phi_i D_ij phi_j|. */
int orc_density(const orc_geom *G, const double *const *tab, int nspin, const double *const *DM,
                double *const *rho, double *const *absrho) {
    const int64_t NN = (int64_t)G->ngrid[0] * G->ngrid[1] * G->ngrid[2];
    int maxno = 0, maxg = 0;
    for (int a = 0; a < G->natom; a++) { if (G->no[a] > maxno) maxno = G->no[a]; if (G->gridn[a] > maxg) maxg = G->gridn[a]; }
    double *ob1 = malloc((size_t)maxno * ORC_BLOCK * sizeof(double));
    double *ob2 = malloc((size_t)maxno * ORC_BLOCK * sizeof(double));
    double *pb = malloc((size_t)maxno * ORC_BLOCK * sizeof(double));
    double *dms = malloc((size_t)maxno * maxno * sizeof(double));
    double *scr = malloc((size_t)(maxg > 0 ? maxg : 1) * sizeof(double));
    double *ascr = malloc((size_t)(maxg > 0 ? maxg : 1) * sizeof(double));
    for (int s = 0; s < nspin; s++) {
        memset(rho[s], 0, (size_t)NN * sizeof(double));
        if (absrho) memset(absrho[s], 0, (size_t)NN * sizeof(double));
    }
    for (int s = 0; s < nspin; s++) {
        for (int64_t q = 0; q < G->npair; q++) {
            int a = G->p_atom[q], b = G->p_natn[q];
            int NO0 = G->no[a], NO1 = G->no[b];
            const double *o1 = tab[a], *o2 = tab[b];
            const int *L1 = G->olg1 + G->olg_off[q], *L2 = G->olg2 + G->olg_off[q];
            int n = G->numolg[q];
            const double *D = DM[s] + G->my_hoff + G->p_hoff[q];
            /* density_matrix_scratch[jst, ist] = DM[dm_index] */
            for (int i = 0; i < NO0; i++) for (int j = 0; j < NO1; j++) dms[(size_t)i * NO1 + j] = D[(int64_t)i * NO1 + j];
            for (int t = 0; t < G->gridn[a]; t++) { scr[t] = 0.0; ascr[t] = 0.0; }
            for (int bs = 0; bs < n; bs += ORC_BLOCK) {
                int pts = n - bs < ORC_BLOCK ? n - bs : ORC_BLOCK;
                for (int k = 0; k < pts; k++) {
                    int t1 = L1[bs + k], t2 = L2[bs + k];
                    for (int i = 0; i < NO0; i++) ob1[(size_t)k * NO0 + i] = o1[(size_t)t1 * NO0 + i];
                    for (int j = 0; j < NO1; j++) ob2[(size_t)k * NO1 + j] = o2[(size_t)t2 * NO1 + j];
                }
                /* product[i, k] = sum_j dms[j, i] * ob2[j, k] */
                for (int k = 0; k < pts; k++)
                    for (int i = 0; i < NO0; i++) {
                        double acc = 0.0;
                        for (int j = 0; j < NO1; j++) acc += dms[(size_t)i * NO1 + j] * ob2[(size_t)k * NO1 + j];
                        pb[(size_t)k * NO0 + i] = acc;
                    }
                for (int k = 0; k < pts; k++) {
                    double value = 0.0;
                    for (int i = 0; i < NO0; i++) value += ob1[(size_t)k * NO0 + i] * pb[(size_t)k * NO0 + i];
                    scr[L1[bs + k]] += value;
                    if (absrho) {
                        double av = 0.0;
                        for (int i = 0; i < NO0; i++)
                            for (int j = 0; j < NO1; j++) av += fabs(ob1[(size_t)k * NO0 + i] * dms[(size_t)i * NO1 + j] * ob2[(size_t)k * NO1 + j]);
                        ascr[L1[bs + k]] += av;
                    }
                }
            }
            const int *gl = G->glist + G->gl_off[a];
            for (int t = 0; t < G->gridn[a]; t++) rho[s][gl[t]] += scr[t];
            if (absrho) for (int t = 0; t < G->gridn[a]; t++) absrho[s][gl[t]] += ascr[t];
        }
    }
    free(ob1); free(ob2); free(pb); free(dms); free(scr); free(ascr);
    return 0;
}

/* _Calc_Force3_Blocked_Kernel! (Force.jl:3346-3392) summed per pair into
 * per-atom sums S[a][xyz] for one channel: S_a = sum_pairs(a,.) sum_p
 * V(g(p)) sum_i dphi_a^xyz[i](p) (D phi_b)(p)_i.  The caller applies the
 * spin weights of Force3_{nospin,spin,nc} (Force.jl:442-444,494-496,557-559).
 * S: natom x 3, row-major, accumulated (+=). */
int orc_force3_sums(const orc_geom *G, const double *const *tab, const double *const *dx,
                    const double *const *dy, const double *const *dz, const double *V,
                    const double *DM, double *S) {
    int maxno = 0;
    for (int a = 0; a < G->natom; a++) if (G->no[a] > maxno) maxno = G->no[a];
    double *ob = malloc((size_t)maxno * ORC_BLOCK * sizeof(double));
    double *pb = malloc((size_t)maxno * ORC_BLOCK * sizeof(double));
    double *dm = malloc((size_t)maxno * maxno * sizeof(double));
    for (int64_t q = 0; q < G->npair; q++) {
        int a = G->p_atom[q], b = G->p_natn[q];
        int NO0 = G->no[a], NO1 = G->no[b];
        const int *L1 = G->olg1 + G->olg_off[q], *L2 = G->olg2 + G->olg_off[q];
        const int *gl = G->glist + G->gl_off[a];
        int n = G->numolg[q];
        const double *D = DM + G->my_hoff + G->p_hoff[q];
        for (int i = 0; i < NO0; i++) for (int j = 0; j < NO1; j++) dm[(size_t)i * NO1 + j] = D[(int64_t)i * NO1 + j];
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int bs = 0; bs < n; bs += ORC_BLOCK) {
            int pts = n - bs < ORC_BLOCK ? n - bs : ORC_BLOCK;
            for (int k = 0; k < pts; k++) {
                int t2 = L2[bs + k];
                for (int j = 0; j < NO1; j++) ob[(size_t)k * NO1 + j] = tab[b][(size_t)t2 * NO1 + j];
            }
            for (int k = 0; k < pts; k++)
                for (int i = 0; i < NO0; i++) {
                    double acc = 0.0;
                    for (int j = 0; j < NO1; j++) acc += dm[(size_t)i * NO1 + j] * ob[(size_t)k * NO1 + j];
                    pb[(size_t)k * NO0 + i] = acc;
                }
            for (int k = 0; k < pts; k++) {
                int t1 = L1[bs + k];
                double pot = V[gl[t1]];
                double lx = 0.0, ly = 0.0, lz = 0.0;
                for (int i = 0; i < NO0; i++) {
                    double pr = pb[(size_t)k * NO0 + i];
                    lx += dx[a][(size_t)t1 * NO0 + i] * pr;
                    ly += dy[a][(size_t)t1 * NO0 + i] * pr;
                    lz += dz[a][(size_t)t1 * NO0 + i] * pr;
                }
                sx += pot * lx; sy += pot * ly; sz += pot * lz;
            }
        }
        S[3 * a] += sx; S[3 * a + 1] += sy; S[3 * a + 2] += sz;
    }
    free(ob); free(pb); free(dm);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Brute force from the set definition, independent of the overlap     */
/* lists and of the block order: for every home grid point g and every */
/* ordered pair of covering sphere entries (eA, eB) at g, the pair is   */
/* (A, B, cell(eA) - cell(eB)).  Used to pin the blocked contractions.  */
/* ------------------------------------------------------------------ */
static int64_t pair_lookup(const orc_geom *G, int a, int b, int cell, const int64_t *first_pair_of_atom) {
    for (int64_t q = first_pair_of_atom[a]; q < G->npair && G->p_atom[q] == a; q++)
        if (G->p_natn[q] == b && G->p_ncn[q] == cell) return q;
    return -1;
}

int orc_bruteforce(const orc_geom *G, const double *const *tab, const double *V, const double *DM,
                   double *H, double *rho, int64_t *n_pairpoints) {
    const int64_t NN = (int64_t)G->ngrid[0] * G->ngrid[1] * G->ngrid[2];
    int C = G->cpy;
    /* entries per grid point */
    int64_t *cnt = calloc((size_t)NN + 1, sizeof(int64_t));
    for (int a = 0; a < G->natom; a++)
        for (int t = 0; t < G->gridn[a]; t++) cnt[G->glist[G->gl_off[a] + t] + 1]++;
    for (int64_t g = 0; g < NN; g++) cnt[g + 1] += cnt[g];
    int64_t tot = cnt[NN];
    int *ea = malloc((size_t)(tot + 1) * sizeof(int)), *et = malloc((size_t)(tot + 1) * sizeof(int));
    int64_t *fill = calloc((size_t)NN, sizeof(int64_t));
    for (int a = 0; a < G->natom; a++)
        for (int t = 0; t < G->gridn[a]; t++) {
            int g = G->glist[G->gl_off[a] + t];
            int64_t pos = cnt[g] + fill[g]++;
            ea[pos] = a; et[pos] = t;
        }
    int64_t *first = malloc((size_t)(G->natom + 1) * sizeof(int64_t));
    for (int a = 0; a <= G->natom; a++) first[a] = G->npair;
    for (int64_t q = G->npair - 1; q >= 0; q--) first[G->p_atom[q]] = q;
    memset(H, 0, (size_t)G->my_hsize * sizeof(double));
    memset(rho, 0, (size_t)NN * sizeof(double));
    int64_t npp = 0;
    for (int64_t g = 0; g < NN; g++) {
        for (int64_t u = cnt[g]; u < cnt[g + 1]; u++)
            for (int64_t w = cnt[g]; w < cnt[g + 1]; w++) {
                int a = ea[u], b = ea[w];
                int ca = G->clist[G->gl_off[a] + et[u]], cb = G->clist[G->gl_off[b] + et[w]];
                int l1 = G->atv_ijk[3 * ca] - G->atv_ijk[3 * cb];
                int l2 = G->atv_ijk[3 * ca + 1] - G->atv_ijk[3 * cb + 1];
                int l3 = G->atv_ijk[3 * ca + 2] - G->atv_ijk[3 * cb + 2];
                if (abs(l1) > C || abs(l2) > C || abs(l3) > C) continue;
                int cell = G->ratv[((l1 + C) * G->nr + (l2 + C)) * G->nr + (l3 + C)];
                int64_t q = pair_lookup(G, a, b, cell, first);
                if (q < 0) continue;  /* pair not in this rank's range */
                npp++;
                int NO0 = G->no[a], NO1 = G->no[b];
                const double *pa = tab[a] + (size_t)et[u] * NO0, *pb = tab[b] + (size_t)et[w] * NO1;
                double v = V[g];
                double *Hb = H + G->p_hoff[q];
                const double *Db = DM + G->my_hoff + G->p_hoff[q];
                double r = 0.0;
                for (int i = 0; i < NO0; i++)
                    for (int j = 0; j < NO1; j++) {
                        Hb[(int64_t)i * NO1 + j] += pa[i] * v * pb[j];
                        r += pa[i] * Db[(int64_t)i * NO1 + j] * pb[j];
                    }
                rho[g] += r;
            }
    }
    for (int64_t h = 0; h < G->my_hsize; h++) H[h] *= G->gridvol;
    *n_pairpoints = npp;
    free(cnt); free(ea); free(et); free(fill); free(first);
    return 0;
}

/* ------------------------------------------------------------------ */
/* CPU baseline: the reference's blocked H + rho assembly with the      */
/* DGEMM of OpenBLAS 0.3.29 (the version the reference pins,            */
/* Manifest.toml:308-311), one BLAS thread per worker (DFT_Setup.jl:368),*/
/* workers on the reference's weighted contiguous pair ranges           */
/* (UCell.jl:50-73).  Timing = max over workers + the rho reduction.    */
/* ------------------------------------------------------------------ */
typedef void (*dgemm_fn)(const char *, const char *, const int64_t *, const int64_t *, const int64_t *,
                         const double *, const double *, const int64_t *, const double *, const int64_t *,
                         const double *, double *, const int64_t *, size_t, size_t);
typedef void (*setthreads_fn)(int);
static dgemm_fn g_dgemm = NULL;

int orc_blas_load(const char *path) {
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return fail(dlerror());
    g_dgemm = (dgemm_fn)dlsym(h, "scipy_dgemm_64_");
    setthreads_fn st = (setthreads_fn)dlsym(h, "scipy_openblas_set_num_threads64_");
    if (!g_dgemm) return fail("scipy_dgemm_64_ not found");
    if (st) st(1);
    return 0;
}

typedef struct {
    const orc_geom *G;
    const double *const *tab;
    const double *V, *DM;
    double *H;        /* my_hsize */
    double *rho;      /* NN, private per worker */
    int64_t q0, q1;   /* pair range */
    double seconds;
    int use_blas;
    int64_t pair_limit;  /* process at most this many pairs (sampling) */
} bl_job;

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void *bl_worker(void *arg) {
    bl_job *J = arg;
    const orc_geom *G = J->G;
    int maxno = 0, maxg = 0;
    for (int a = 0; a < G->natom; a++) { if (G->no[a] > maxno) maxno = G->no[a]; if (G->gridn[a] > maxg) maxg = G->gridn[a]; }
    double *ob1 = malloc((size_t)maxno * ORC_BLOCK * sizeof(double));
    double *ob2 = malloc((size_t)maxno * ORC_BLOCK * sizeof(double));
    double *wb = malloc((size_t)maxno * ORC_BLOCK * sizeof(double));
    double *ps = malloc((size_t)maxno * maxno * sizeof(double));
    double *dms = malloc((size_t)maxno * maxno * sizeof(double));
    double *scr = malloc((size_t)(maxg + 1) * sizeof(double));
    double t0 = now_s();
    int64_t done = 0;
    for (int64_t q = J->q0; q < J->q1 && done < J->pair_limit; q++, done++) {
        int a = G->p_atom[q], b = G->p_natn[q];
        int64_t NO0 = G->no[a], NO1 = G->no[b];
        const double *o1 = J->tab[a], *o2 = J->tab[b];
        const int *L1 = G->olg1 + G->olg_off[q], *L2 = G->olg2 + G->olg_off[q];
        const int *gl = G->glist + G->gl_off[a];
        int n = G->numolg[q];
        double *Hb = J->H + G->p_hoff[q];
        const double *D = J->DM + G->my_hoff + G->p_hoff[q];
        for (int64_t i = 0; i < NO0 * NO1; i++) { Hb[i] = 0.0; dms[i] = D[i]; }
        for (int t = 0; t < G->gridn[a]; t++) scr[t] = 0.0;
        for (int bs = 0; bs < n; bs += ORC_BLOCK) {
            int64_t pts = n - bs < ORC_BLOCK ? n - bs : ORC_BLOCK;
            for (int64_t k = 0; k < pts; k++) {
                int t1 = L1[bs + k], t2 = L2[bs + k];
                double v = J->V[gl[t1]];
                for (int64_t i = 0; i < NO0; i++) { double x = o1[(size_t)t1 * NO0 + i]; ob1[k * NO0 + i] = x; wb[k * NO0 + i] = x * v; }
                for (int64_t j = 0; j < NO1; j++) ob2[k * NO1 + j] = o2[(size_t)t2 * NO1 + j];
            }
            const double one = 1.0, zero = 0.0;
            /* H: ps[NO1 x NO0] = ob2[NO1 x pts] * wb[NO0 x pts]^T */
            if (J->use_blas) g_dgemm("N", "T", &NO1, &NO0, &pts, &one, ob2, &NO1, wb, &NO0, &zero, ps, &NO1, 1, 1);
            else for (int64_t i = 0; i < NO0; i++) for (int64_t j = 0; j < NO1; j++) { double s = 0; for (int64_t k = 0; k < pts; k++) s += ob2[k * NO1 + j] * wb[k * NO0 + i]; ps[i * NO1 + j] = s; }
            for (int64_t i = 0; i < NO0 * NO1; i++) Hb[i] += ps[i];
            /* rho: wb[NO0 x pts] = dms^T[NO0 x NO1] * ob2[NO1 x pts]  (dms stored [i*NO1+j] = column-major NO1 x NO0) */
            if (J->use_blas) g_dgemm("T", "N", &NO0, &pts, &NO1, &one, dms, &NO1, ob2, &NO1, &zero, wb, &NO0, 1, 1);
            else for (int64_t k = 0; k < pts; k++) for (int64_t i = 0; i < NO0; i++) { double s = 0; for (int64_t j = 0; j < NO1; j++) s += dms[i * NO1 + j] * ob2[k * NO1 + j]; wb[k * NO0 + i] = s; }
            for (int64_t k = 0; k < pts; k++) {
                double value = 0.0;
                for (int64_t i = 0; i < NO0; i++) value += ob1[k * NO0 + i] * wb[k * NO0 + i];
                scr[L1[bs + k]] += value;
            }
        }
        for (int64_t i = 0; i < NO0 * NO1; i++) Hb[i] *= G->gridvol;
        for (int t = 0; t < G->gridn[a]; t++) J->rho[gl[t]] += scr[t];
    }
    J->seconds = now_s() - t0;
    free(ob1); free(ob2); free(wb); free(ps); free(dms); free(scr);
    return NULL;
}

/* Runs the baseline with `nworkers` threads on the reference's weighted
 * contiguous split of the first `pair_limit_total` pairs (all pairs when
 * <= 0).  Returns wall seconds (max worker + rho reduction) in *seconds and
 * the number of pairs processed in *pairs_done. */
int orc_cpu_baseline(const orc_geom *G, const double *const *tab, const double *V, const double *DM,
                     double *H, double *rho, int nworkers, int use_blas, int64_t pair_limit_total,
                     double *seconds, int64_t *pairs_done) {
    if (use_blas && !g_dgemm) return fail("BLAS not loaded");
    const int64_t NN = (int64_t)G->ngrid[0] * G->ngrid[1] * G->ngrid[2];
    int64_t npair = G->npair;
    if (pair_limit_total > 0 && pair_limit_total < npair) npair = pair_limit_total;
    if (nworkers < 1) nworkers = 1;
    if (nworkers > npair) nworkers = (int)npair;
    /* weighted contiguous ranges over the sampled pairs, weights as UCell.jl:95-97 */
    double *cum = malloc((size_t)npair * sizeof(double));
    double acc = 0;
    for (int64_t q = 0; q < npair; q++) {
        int a = G->p_atom[q], b = G->p_natn[q];
        double w = (double)G->no[a] * G->no[b] * (double)(G->gridn[a] < G->gridn[b] ? G->gridn[a] : G->gridn[b]);
        acc += w < 1 ? 1 : w;
        cum[q] = acc;
    }
    bl_job *J = calloc((size_t)nworkers, sizeof(bl_job));
    pthread_t *th = calloc((size_t)nworkers, sizeof(pthread_t));
    int64_t beg = 0;
    for (int w = 0; w < nworkers; w++) {
        int64_t end = npair;
        if (w < nworkers - 1) {
            double target = acc * (w + 1) / nworkers;
            end = beg + 1;
            while (end < npair - (nworkers - 1 - w) && cum[end - 1] < target) end++;
        }
        J[w] = (bl_job){G, tab, V, DM, H, calloc((size_t)NN, sizeof(double)), beg, end, 0.0, use_blas, INT64_MAX};
        beg = end;
    }
    double t0 = now_s();
    for (int w = 0; w < nworkers; w++) pthread_create(&th[w], NULL, bl_worker, &J[w]);
    for (int w = 0; w < nworkers; w++) pthread_join(th[w], NULL);
    double t1 = now_s();
    memset(rho, 0, (size_t)NN * sizeof(double));
    for (int w = 0; w < nworkers; w++) { for (int64_t g = 0; g < NN; g++) rho[g] += J[w].rho[g]; free(J[w].rho); }
    double t2 = now_s();
    *seconds = (t1 - t0) + (t2 - t1);
    *pairs_done = npair;
    free(cum); free(J); free(th);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Atomic densities on the grid: Set_AdenPCC_Grid.jl:1-142            */
/* ------------------------------------------------------------------ */

/* KumoF, utils/KumoF.jl:1-62.  xv, rv: the log mesh (N points, 0-based);
 * yv: N + 4 values with mesh point k at yv[k + 1] (Julia's yv[k + 2]). */
double orc_kumof(int N, double x, const double *xv, const double *rv, const double *yv) {
    if (x < xv[0]) {
        /* KumoF.jl:3-40, m = 5 (1-based): rv[m] -> rv[4], yv[m] -> yv[4] */
        double r = exp(x);
        double rm = rv[4];
        double h1 = rv[3] - rv[2];
        double h2 = rv[4] - rv[3];
        double h3 = rv[5] - rv[4];
        double f1 = yv[2], f2 = yv[3], f3 = yv[4], f4 = yv[5];
        double g1 = ((f3 - f2) * h1 / h2 + (f2 - f1) * h2 / h1) / (h1 + h2);
        double g2 = ((f4 - f3) * h2 / h3 + (f3 - f2) * h3 / h2) / (h2 + h3);
        double x1 = rm - rv[3];
        double x2 = rm - rv[4];
        double y1 = x1 / h2;
        double y2 = x2 / h2;
        double y12 = y1 * y1;
        double y22 = y2 * y2;
        double A = 2 * f2 + h2 * g1;
        double B = 2 * f3 - h2 * g2;
        double C = 3 * f2 + h2 * g1;
        double D = 3 * f3 - h2 * g2;
        double E = C + A * y2;
        double F = D - B * y1;
        double f = y22 * E + y12 * F;
        double df = (2 * y2 / h2) * E + y22 * A / h2 + (2 * y1 / h2) * F - y12 * B / h2;
        double a = 0.5 * df / rm;
        double b = f - a * rm * rm;
        return a * r * r + b;
    }
    /* KumoF.jl:41-60: 1-based i = floor(tmp) + 1 reads yv[i .. i+3] */
    double xmin = xv[0], xmax = xv[N - 1];
    if (x > xmax) x = xmax;
    if (x < xmin) x = xmin;
    double tmp = (double)(N - 1) * (x - xmin) / (xmax - xmin);
    int i = (int)floor(tmp);
    double dt = tmp - i;
    double d0 = yv[i + 1] - yv[i];
    double d1 = yv[i + 2] - yv[i + 1];
    double d2 = yv[i + 3] - yv[i + 2];
    double a = d0 - 2 * d1 + d2;
    double b = -2 * d0 + 3 * d1 - d2;
    double c = d0 + d1;
    return yv[i + 1] + 0.5 * (((a * dt + b) * dt + c) * dt);
}

/* Set_AdenPCC_Grid (Set_AdenPCC_Grid.jl:1-142) with one rank: the atom loop
 * (:56-115) over every atom's grid list, then the nc rotation (:125-138).
 * Per species s: npao[s], pxv/prv[s] (npao), pden[s] (npao + 4); nvps[s],
 * vxv/vrv[s] (nvps), vpcc[s] (nvps + 4).  init_nspin: natom x 2;
 * init_angle: natom x 2 (theta, phi) or NULL.  Outputs NN each. */
int orc_aden_pcc(const orc_geom *G, int nspin, const int *atom2spe, const int *npao, const double *const *pxv,
                 const double *const *prv, const double *const *pden, const int *nvps, const double *const *vxv,
                 const double *const *vrv, const double *const *vpcc, const double *init_nspin,
                 const double *init_angle, double *aden, double *pcc, double *const *rho) {
    if (nspin != 1 && nspin != 2 && nspin != 4) return fail("nspin must be 1, 2 or 4");
    const int N1 = G->ngrid[0], N2 = G->ngrid[1], N3 = G->ngrid[2];
    const int64_t NN = (int64_t)N1 * N2 * N3;
    double glat[9];
    for (int j = 0; j < 3; j++) { glat[j] = G->lat[j] / N1; glat[3 + j] = G->lat[3 + j] / N2; glat[6 + j] = G->lat[6 + j] / N3; }
    memset(aden, 0, sizeof(double) * NN);
    memset(pcc, 0, sizeof(double) * NN);
    for (int s = 0; s < nspin; s++) memset(rho[s], 0, sizeof(double) * NN);
    for (int atom = 0; atom < G->natom; atom++) {
        const int spe = atom2spe[atom];
        const double Nu = init_nspin[2 * atom], Nd = init_nspin[2 * atom + 1];
        const double Nele = Nu + Nd;
        const double ocupcy_u = Nele > 1e-15 ? Nu / Nele : 0.0;
        const double ocupcy_p = Nele > 1e-15 ? Nd / Nele : 0.0;
        const double theta = init_angle ? init_angle[2 * atom] : 0.0;
        const double phi = init_angle ? init_angle[2 * atom + 1] : 0.0;
        for (int xyz = 0; xyz < G->gridn[atom]; xyz++) {
            int GNc = G->glist[G->gl_off[atom] + xyz];
            int GRc = G->clist[G->gl_off[atom] + xyz];
            int n1 = GNc / (N2 * N3);
            int n2 = (GNc - n1 * N2 * N3) / N3;
            int n3 = GNc - n1 * N2 * N3 - n2 * N3;
            double Cx = n1 * glat[0] + n2 * glat[3] + n3 * glat[6] + G->origin[0];
            double Cy = n1 * glat[1] + n2 * glat[4] + n3 * glat[7] + G->origin[1];
            double Cz = n1 * glat[2] + n2 * glat[5] + n3 * glat[8] + G->origin[2];
            double dx = Cx + G->atv[3 * GRc] - G->gxyz[3 * atom];
            double dy = Cy + G->atv[3 * GRc + 1] - G->gxyz[3 * atom + 1];
            double dz = Cz + G->atv[3 * GRc + 2] - G->gxyz[3 * atom + 2];
            double x = 0.5 * log(dx * dx + dy * dy + dz * dz);
            double ADen = orc_kumof(npao[spe], x, pxv[spe], prv[spe], pden[spe]);
            double PCCDen = orc_kumof(nvps[spe], x, vxv[spe], vrv[spe], vpcc[spe]);
            aden[GNc] += 0.5 * ADen;
            pcc[GNc] += 0.5 * PCCDen;
            if (nspin == 1) {
                rho[0][GNc] += 0.5 * ADen;
            } else if (nspin == 2) {
                rho[0][GNc] += ocupcy_u * ADen;
                rho[1][GNc] += ocupcy_p * ADen;
            } else {
                double mag = (ocupcy_u - ocupcy_p) * ADen;
                rho[0][GNc] += ADen;
                rho[1][GNc] += mag * sin(theta) * cos(phi);
                rho[2][GNc] += mag * sin(theta) * sin(phi);
                rho[3][GNc] += mag * cos(theta);
            }
        }
    }
    if (nspin == 4)
        for (int64_t i = 0; i < NN; i++) {
            double r = rho[0][i], mx = rho[1][i], my = rho[2][i], mz = rho[3][i];
            rho[0][i] = 0.5 * (r + mz);
            rho[1][i] = 0.5 * (r - mz);
            rho[2][i] = 0.5 * mx;
            rho[3][i] = -0.5 * my;
        }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Poisson: Poisson.jl:1-108                                          */
/* ------------------------------------------------------------------ */

/* 1-D DFT of n complex values in place, sign -1 forward (FFTW's fft),
 * +1 backward scaled by 1/n (FFTW's normalised ifft): recursive mixed-radix
 * Cooley-Tukey on the smallest prime factor, direct sums for a prime n. */
static void dft_rec(int n, int sign, const double *in, int stride, double *out, double *tmp) {
    if (n == 1) { out[0] = in[0]; out[1] = in[1]; return; }
    int p = 2;
    while (p * p <= n && n % p) p++;
    if (n % p) p = n;
    const int m = n / p;
    if (p == n) {
        for (int k = 0; k < n; k++) {
            double sr = 0.0, si = 0.0;
            for (int j = 0; j < n; j++) {
                double ang = sign * 2.0 * ORC_PI * (double)((int64_t)j * k % n) / n;
                double c = cos(ang), s = sin(ang);
                double xr = in[2 * (int64_t)j * stride], xi = in[2 * (int64_t)j * stride + 1];
                sr += xr * c - xi * s;
                si += xr * s + xi * c;
            }
            out[2 * k] = sr;
            out[2 * k + 1] = si;
        }
        return;
    }
    /* p sub-transforms of length m over the residues r: x[r + p j] */
    for (int r = 0; r < p; r++) dft_rec(m, sign, in + 2 * (int64_t)r * stride, stride * p, tmp + 2 * (int64_t)r * m, out);
    for (int k = 0; k < m; k++)
        for (int q = 0; q < p; q++) {
            const int kk = k + m * q;
            double sr = 0.0, si = 0.0;
            for (int r = 0; r < p; r++) {
                double ang = sign * 2.0 * ORC_PI * (double)((int64_t)r * kk % n) / n;
                double c = cos(ang), s = sin(ang);
                double yr = tmp[2 * ((int64_t)r * m + k)], yi = tmp[2 * ((int64_t)r * m + k) + 1];
                sr += yr * c - yi * s;
                si += yr * s + yi * c;
            }
            out[2 * kk] = sr;
            out[2 * kk + 1] = si;
        }
}

static void dft_line(int n, int sign, double *line, double *work) {
    /* work: 6 n doubles (result, scratch of the recursion, copy of the input) */
    double *res = work, *scr = work + 2 * n, *cp = work + 4 * n;
    memcpy(cp, line, sizeof(double) * 2 * n);
    dft_rec(n, sign, cp, 1, res, scr);
    for (int k = 0; k < 2 * n; k++) line[k] = sign > 0 ? res[k] / n : res[k];
}

/* transform along one axis of the complex N1 x N2 x N3 array (n3 fastest) */
static void dft_axis(double *z, const int *N, int axis, int sign) {
    const int n = N[axis];
    double *line = malloc(sizeof(double) * 2 * n), *work = malloc(sizeof(double) * 6 * n);
    const int64_t s3 = 1, s2 = N[2], s1 = (int64_t)N[1] * N[2];
    const int64_t st = axis == 0 ? s1 : (axis == 1 ? s2 : s3);
    const int o1 = axis == 0 ? 1 : N[0], o2 = axis == 1 ? 1 : N[1], o3 = axis == 2 ? 1 : N[2];
    for (int a = 0; a < o1; a++)
        for (int b = 0; b < o2; b++)
            for (int c = 0; c < o3; c++) {
                const int64_t base = a * s1 + b * s2 + c * s3;
                for (int j = 0; j < n; j++) { line[2 * j] = z[2 * (base + j * st)]; line[2 * j + 1] = z[2 * (base + j * st) + 1]; }
                dft_line(n, sign, line, work);
                for (int j = 0; j < n; j++) { z[2 * (base + j * st)] = line[2 * j]; z[2 * (base + j * st) + 1] = line[2 * j + 1]; }
            }
    free(line);
    free(work);
}

/* Solve_Poisson!(Ngrid, Recvecs, Density_xyz, dVHart_Grid), Poisson.jl:29-108:
 * forward 1-D transforms along axes 3, 2, 1; times 4 pi / |G|^2 with the
 * signed indices of :65-80; G = 0 set to zero; normalised inverse transforms
 * along axes 1, 2, 3; the real part.  rec: Recvecs column-major
 * (rec[3 j + i] = Recvecs[i+1, j+1]). */
int orc_poisson(const int *ngrid, const double *rec, const double *density_xyz, double *vhart) {
    const int N1 = ngrid[0], N2 = ngrid[1], N3 = ngrid[2];
    const int64_t NN = (int64_t)N1 * N2 * N3;
    double *z = malloc(sizeof(double) * 2 * NN);
    if (!z) return fail("oom");
    for (int64_t g = 0; g < NN; g++) { z[2 * g] = density_xyz[g]; z[2 * g + 1] = 0.0; }
    dft_axis(z, ngrid, 2, -1);
    dft_axis(z, ngrid, 1, -1);
    dft_axis(z, ngrid, 0, -1);
    const double temp = 4 * ORC_PI;
    const int Ngrid11 = N1 / 2, Ngrid22 = N2 / 2, Ngrid33 = N3 / 2;
    for (int i = 0; i < N1; i++)
        for (int j = 0; j < N2; j++)
            for (int k = 0; k < N3; k++) {
                int sk1 = i < Ngrid11 ? i : i - N1;
                int sk2 = j < Ngrid22 ? j : j - N2;
                int sk3 = k < Ngrid33 ? k : k - N3;
                double Gx = sk1 * rec[0] + sk2 * rec[1] + sk3 * rec[2];
                double Gy = sk1 * rec[3] + sk2 * rec[4] + sk3 * rec[5];
                double Gz = sk1 * rec[6] + sk2 * rec[7] + sk3 * rec[8];
                double g2 = Gx * Gx + Gy * Gy + Gz * Gz;
                int64_t g = ((int64_t)i * N2 + j) * N3 + k;
                z[2 * g] = temp * z[2 * g] / g2;
                z[2 * g + 1] = temp * z[2 * g + 1] / g2;
            }
    z[0] = 0.0;
    z[1] = 0.0;
    dft_axis(z, ngrid, 0, 1);
    dft_axis(z, ngrid, 1, 1);
    dft_axis(z, ngrid, 2, 1);
    for (int64_t g = 0; g < NN; g++) vhart[g] = z[2 * g];
    free(z);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Either side of the diagonalisation                                 */
/* ------------------------------------------------------------------ */

/* cispi(x) = (cos(pi x), sin(pi x)) with the argument reduced exactly as
 * Julia's sincospi does for the phase arguments 2 k.R (reduced mod 2 first) */
static void orc_cispi(double x, double *c, double *s) {
    double r = fmod(x, 2.0);
    if (r < 0) r += 2.0;
    *c = cos(ORC_PI * r);
    *s = sin(ORC_PI * r);
}

static int64_t *orbital_offsets(const orc_geom *G) {
    int64_t *mp = malloc(sizeof(int64_t) * (G->natom + 1));
    mp[0] = 0;
    for (int a = 0; a < G->natom; a++) mp[a + 1] = mp[a] + G->no[a];
    return mp;
}

/* HS_matrix! with k (HS_matrix.jl:88-110, 143-166) over the geometry's local
 * pairs (all pairs for one rank): dense[m] (fsize x fsize, column-major,
 * interleaved complex) = sum of flat[m][pair][ist, jst] * cispi(2 kRn). */
int orc_hs_matrix(const orc_geom *G, const double *kpt, int nmat, const double *const *flat, double *const *dense) {
    int64_t *mp = orbital_offsets(G);
    const int64_t fsize = mp[G->natom];
    for (int m = 0; m < nmat; m++) memset(dense[m], 0, sizeof(double) * 2 * fsize * fsize);
    const double ka = kpt[0], kb = kpt[1], kc = kpt[2];
    for (int64_t p = 0; p < G->npair; p++) {
        const int atom = G->p_atom[p], jatom = G->p_natn[p], cell = G->p_ncn[p];
        const int NO0 = G->no[atom], NO1 = G->no[jatom];
        const int64_t Anum = mp[atom], Bnum = mp[jatom];
        const double kRn = ka * G->atv_ijk[3 * cell] + kb * G->atv_ijk[3 * cell + 1] + kc * G->atv_ijk[3 * cell + 2];
        double c, s;
        orc_cispi(2 * kRn, &c, &s);
        int64_t hst = G->p_hoff[p];
        for (int ist = 0; ist < NO0; ist++)
            for (int jst = 0; jst < NO1; jst++, hst++) {
                const int64_t o = (Anum + ist) + (Bnum + jst) * fsize;
                for (int m = 0; m < nmat; m++) {
                    dense[m][2 * o] += flat[m][hst] * c;
                    dense[m][2 * o + 1] += flat[m][hst] * s;
                }
            }
    }
    free(mp);
    return 0;
}

/* HS_matrix_NC! with k and flat Hks (HS_matrix.jl:211-257): H ((2 fsize)^2,
 * column-major, interleaved complex) is zeroed, then for each pair and
 * (ist, jst), with ex = cispi(2 kRn):
 *   H[A+i, B+j]             += (Hks[0] + im*iHks[0]) * ex
 *   H[fsize+A+i, fsize+B+j] += (Hks[1] + im*iHks[1]) * ex
 *   H[A+i, fsize+B+j]       += (Hks[2] + im*(Hks[3] + iHks[2])) * ex
 * iHks are the nested arrays flattened in (atom, Rn, ist, jst) order. */
static void cmul_add(double *z, double ar, double ai, double br, double bi) {
    z[0] += ar * br - ai * bi;
    z[1] += ar * bi + ai * br;
}

int orc_hs_matrix_nc(const orc_geom *G, const double *kpt, const double *const *hks, const double *const *ihks,
                     double *H) {
    int64_t *mp = orbital_offsets(G);
    const int64_t fsize = mp[G->natom], ld = 2 * fsize;
    memset(H, 0, sizeof(double) * 2 * ld * ld);
    const double ka = kpt[0], kb = kpt[1], kc = kpt[2];
    for (int64_t p = 0; p < G->npair; p++) {
        const int atom = G->p_atom[p], jatom = G->p_natn[p], cell = G->p_ncn[p];
        const int NO0 = G->no[atom], NO1 = G->no[jatom];
        const int64_t Anum = mp[atom], Bnum = mp[jatom];
        const double kRn = ka * G->atv_ijk[3 * cell] + kb * G->atv_ijk[3 * cell + 1] + kc * G->atv_ijk[3 * cell + 2];
        double c, s;
        orc_cispi(2 * kRn, &c, &s);
        int64_t hst = G->p_hoff[p];
        for (int ist = 0; ist < NO0; ist++)
            for (int jst = 0; jst < NO1; jst++, hst++) {
                const int64_t uu = (Anum + ist) + (Bnum + jst) * ld;
                const int64_t dd = (fsize + Anum + ist) + (fsize + Bnum + jst) * ld;
                const int64_t ud = (Anum + ist) + (fsize + Bnum + jst) * ld;
                cmul_add(H + 2 * uu, hks[0][hst], ihks[0][hst], c, s);
                cmul_add(H + 2 * dd, hks[1][hst], ihks[1][hst], c, s);
                cmul_add(H + 2 * ud, hks[2][hst], hks[3][hst] + ihks[2][hst], c, s);
            }
    }
    free(mp);
    return 0;
}

/* Calc_DM_Crystal_NonCollinear! (Crystal_DFT.jl:342-410) and the k-point sums
 * of Calc_iDM_Crystal_NonCollinear! (:413-486): per k-point the dense
 * DM_dense = C C^H ((2 fsize)^2, a plain triple loop), then for each pair
 * with ex = cispi(2 kRn) * scale:
 *   DM[0] += Re(D[B+j, A+i] ex), DM[1] += Re(D[f+B+j, f+A+i] ex),
 *   DM[2] += Re(D[f+B+j, A+i] ex), DM[3] += Im(D[f+B+j, A+i] ex),
 *   iDM[0] += Im(D[B+j, A+i] ex), iDM[1] += Im(D[f+B+j, f+A+i] ex)
 * (iDM may be NULL).  C[k]: (2 fsize) x nstates interleaved complex. */
int orc_dm_nc(const orc_geom *G, int nk, const double *kpts, int64_t nstates, const double *const *C, double scale,
              double *const *DM, double *const *iDM) {
    int64_t *mp = orbital_offsets(G);
    const int64_t fsize = mp[G->natom], ld = 2 * fsize;
    double *D = malloc(sizeof(double) * 2 * ld * ld);
    if (!D) { free(mp); return fail("oom"); }
    for (int m = 0; m < 4; m++) memset(DM[m], 0, sizeof(double) * G->my_hsize);
    if (iDM)
        for (int m = 0; m < 2; m++) memset(iDM[m], 0, sizeof(double) * G->my_hsize);
    for (int k = 0; k < nk; k++) {
        const double *c = C[k];
        for (int64_t col = 0; col < ld; col++)
            for (int64_t row = 0; row < ld; row++) {
                double re = 0.0, im = 0.0;
                for (int64_t mu = 0; mu < nstates; mu++) {
                    const double ar = c[2 * (row + mu * ld)], ai = c[2 * (row + mu * ld) + 1];
                    const double br = c[2 * (col + mu * ld)], bi = c[2 * (col + mu * ld) + 1];
                    re += ar * br + ai * bi; /* a * conj(b) */
                    im += ai * br - ar * bi;
                }
                D[2 * (row + col * ld)] = re;
                D[2 * (row + col * ld) + 1] = im;
            }
        const double ka = kpts[3 * k], kb = kpts[3 * k + 1], kc = kpts[3 * k + 2];
        for (int64_t p = 0; p < G->npair; p++) {
            const int atom = G->p_atom[p], jatom = G->p_natn[p], cell = G->p_ncn[p];
            const int NO0 = G->no[atom], NO1 = G->no[jatom];
            const int64_t Anum = mp[atom], Bnum = mp[jatom];
            const double kRn = ka * G->atv_ijk[3 * cell] + kb * G->atv_ijk[3 * cell + 1] + kc * G->atv_ijk[3 * cell + 2];
            double er, ei;
            orc_cispi(2 * kRn, &er, &ei);
            er = er * scale;
            ei = ei * scale;
            int64_t DMst = G->p_hoff[p];
            for (int ist = 0; ist < NO0; ist++)
                for (int jst = 0; jst < NO1; jst++, DMst++) {
                    const double *uu = D + 2 * ((Bnum + jst) + (Anum + ist) * ld);
                    const double *dd = D + 2 * ((fsize + Bnum + jst) + (fsize + Anum + ist) * ld);
                    const double *ud = D + 2 * ((fsize + Bnum + jst) + (Anum + ist) * ld);
                    DM[0][DMst] += uu[0] * er - uu[1] * ei;
                    DM[1][DMst] += dd[0] * er - dd[1] * ei;
                    DM[2][DMst] += ud[0] * er - ud[1] * ei;
                    DM[3][DMst] += ud[0] * ei + ud[1] * er;
                    if (iDM) {
                        iDM[0][DMst] += uu[0] * ei + uu[1] * er;
                        iDM[1][DMst] += dd[0] * ei + dd[1] * er;
                    }
                }
        }
    }
    free(D);
    free(mp);
    return 0;
}

/* Calc_DM_Crystal_Collinear_nopol! (Crystal_DFT.jl:219-270): per k-point the
 * dense DM_dense = C C^H (mul!, a plain triple loop here), then for each pair
 * DM[DMst] += real(DM_dense[Bnum+jst, Anum+ist] * cispi(2 kRn) * scale).
 * C[k]: fsize x nstates interleaved complex, column-major. */
int orc_dm_collinear(const orc_geom *G, int nk, const double *kpts, int64_t nstates, const double *const *C,
                     double scale, double *DM) {
    int64_t *mp = orbital_offsets(G);
    const int64_t fsize = mp[G->natom];
    double *D = malloc(sizeof(double) * 2 * fsize * fsize);
    if (!D) { free(mp); return fail("oom"); }
    memset(DM, 0, sizeof(double) * G->my_hsize);
    for (int k = 0; k < nk; k++) {
        const double *c = C[k];
        for (int64_t col = 0; col < fsize; col++)
            for (int64_t row = 0; row < fsize; row++) {
                double re = 0.0, im = 0.0;
                for (int64_t mu = 0; mu < nstates; mu++) {
                    const double ar = c[2 * (row + mu * fsize)], ai = c[2 * (row + mu * fsize) + 1];
                    const double br = c[2 * (col + mu * fsize)], bi = c[2 * (col + mu * fsize) + 1];
                    re += ar * br + ai * bi;  /* a * conj(b) */
                    im += ai * br - ar * bi;
                }
                D[2 * (row + col * fsize)] = re;
                D[2 * (row + col * fsize) + 1] = im;
            }
        const double ka = kpts[3 * k], kb = kpts[3 * k + 1], kc = kpts[3 * k + 2];
        for (int64_t p = 0; p < G->npair; p++) {
            const int atom = G->p_atom[p], jatom = G->p_natn[p], cell = G->p_ncn[p];
            const int NO0 = G->no[atom], NO1 = G->no[jatom];
            const int64_t Anum = mp[atom], Bnum = mp[jatom];
            const double kRn = ka * G->atv_ijk[3 * cell] + kb * G->atv_ijk[3 * cell + 1] + kc * G->atv_ijk[3 * cell + 2];
            double er, ei;
            orc_cispi(2 * kRn, &er, &ei);
            er = er * scale;
            ei = ei * scale;
            int64_t DMst = G->p_hoff[p];
            for (int ist = 0; ist < NO0; ist++)
                for (int jst = 0; jst < NO1; jst++, DMst++) {
                    const int64_t o = (Bnum + jst) + (Anum + ist) * fsize;
                    DM[DMst] += D[2 * o] * er - D[2 * o + 1] * ei;
                }
        }
    }
    free(D);
    free(mp);
    return 0;
}
