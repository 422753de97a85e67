/*
 * lcpao_grid.h -- C ABI of the MI355X (gfx950) real-space grid library.
 *
 * The library computes, for an LCPAO basis on a real-space grid,
 *   H_{Ai,Bj,Rn} = GridVol * sum_p phi_Ai(p) V(g(p)) phi_Bj(p)     (grid Hamiltonian)
 *   rho(g)       = sum_pairs sum_{p: g(p)=g} phi_A(p)^T D_pair phi_B(p)  (density)
 *   S_A          = sum_pairs(A,.) sum_p V(g(p)) dphi_A(p)^T (D_pair phi_B(p)) (Force3)
 * over the atom-pair overlap lists of the host code.  Every entry point is a
 * drop-in for one routine of LCPAODFT.jl; the citation on each declaration
 * names the reference routine (file:line, paths relative to the reference
 * repository root) whose argument meaning and output semantics it keeps.
 *
 * Conventions
 *  - All functions return 0 on success and a non-zero LCPAO_E* code on
 *    failure; lcpao_grid_last_error() then describes the failure.  No C++
 *    exception crosses this boundary.
 *  - Host pointers are read during the call only; nothing is retained.
 *  - Pointers named d_* are device pointers (hipMalloc'd on the context's
 *    device); they let a caller keep V, D, H and rho resident in HBM.
 *  - A context is bound to one device and one stream and is not
 *    thread-safe.  Calls are synchronous on return.
 *  - Index arrays use the reference's values: grid indices, cell indices and
 *    overlap indices are 0-based; atom indices use geom->atom_base (1 for the
 *    Julia arrays MPI_atom / MPI_natn, 0 for C callers).
 */
#ifndef LCPAO_GRID_H
#define LCPAO_GRID_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    LCPAO_OK = 0,
    LCPAO_EINVAL = 1,    /* bad argument / inconsistent geometry */
    LCPAO_ERANGE = 2,    /* size beyond a supported maximum (more than 144 orbitals on an
                            atom, the largest basis Read_PAO's basis string describes) */
    LCPAO_EDEVICE = 3,   /* HIP runtime error, no device, or out of device memory */
    LCPAO_ESTATE = 4,    /* call order (e.g. contraction before tables are set) */
    LCPAO_EMESH = 5      /* radial mesh index out of range in table generation */
};

enum {
    LCPAO_FP64 = 0,      /* plain FP64, the parity form */
    LCPAO_FP32 = 1,      /* orbital products in FP32, pair partials summed in FP64 */
    LCPAO_BF16X3 = 2     /* 3-term split-BF16 MFMA, FP32 accumulation, FP64 pair sums */
};

typedef struct lcpao_ctx lcpao_ctx;

/* The rank-local part of the reference's UCell (UCell.jl:1-46, 210-294).
 * Arrays indexed by atom have natom entries; arrays indexed by pair have
 * npairs entries, in the rank-local pair order of split_system_grid
 * (UCell.jl:76-132). */
typedef struct {
    int32_t natom;
    int32_t atom_base;               /* 0 or 1: base of pair_atom / pair_natn values */
    int32_t ngrid[3];                /* Ngrid1, Ngrid2, Ngrid3 (g = n1*N2*N3 + n2*N3 + n3) */
    int32_t cpycell;                 /* CpyCell: cell index <-> (i,j,k) of Generation_ATV! */
    double gridvol;                  /* GridVol */
    const int32_t *no;               /* Total_NumOrbs[natom] */
    const int32_t *gridn;            /* GridN_Atom[natom] */
    const int32_t *const *grid_list; /* GridListAtom[atom][GridN_Atom[atom]] */
    const int32_t *const *cell_list; /* CellListAtom[atom][...]; may be NULL unless tables are built on device */
    int64_t npairs;                  /* MPI_size */
    const int32_t *pair_atom;        /* MPI_atom[npairs] */
    const int32_t *pair_natn;        /* MPI_natn[npairs] */
    const int32_t *pair_ncn;         /* MPI_ncn[npairs] */
    const int32_t *numolg;           /* MPI_NumOLG[npairs] */
    const int32_t *const *olg1;      /* MPI_GListTAtoms1[pair][NumOLG] */
    const int32_t *const *olg2;      /* MPI_GListTAtoms2[pair][NumOLG] */
    /* Needed only for on-device table generation (may be NULL otherwise). */
    const double *latvecs;           /* 3x3 row-major, rows are lattice vectors */
    const double *origin;            /* Grid_Origin[3] */
    const double *gxyz;              /* natom x 3 Cartesian positions (bohr) */
    const int32_t *atom2spe;         /* natom, 0-based species */
} lcpao_geom;

/* Radial PAO data of one species (Read_PAO.jl:1-18): the mesh RV and the
 * radial functions RWF[l][p] for l <= maxl, p < num_basis[l]. */
typedef struct {
    int32_t maxl;                    /* Spe_MaxL_Basis (<= 3) */
    int32_t num_basis[4];            /* Spe_Num_Basis */
    int32_t nmesh;                   /* Spe_Num_Mesh_PAO */
    const double *rv;                /* Spe_PAO_RV[nmesh] */
    const double *const *rwf;        /* rwf[l*16 + p] -> Spe_PAO_RWF[l+1][p+1][nmesh] (64 entries, p < 16) */
} lcpao_pao;

/* Builds a context for one geometry: copies the lists to the device and
 * prepares the work decomposition.  device < 0 selects the current device.
 * Replaces the per-call setup inside Set_Hamiltonian! / Set_Density_Grid! /
 * Force3 (Hamiltonian.jl:235-277, Set_Density_Grid.jl:60-109,
 * Force.jl:407-451), which read the same UCell fields. */
int lcpao_grid_create(lcpao_ctx **out, const lcpao_geom *geom, int device);

/* Uploads the orbital tables Orbs_Grid.data[atom] (NO_A x GridN_A,
 * column-major, orbital fastest) of PackedOrbitalsGrid
 * (Set_Orbitals_Grid.jl:1-42). */
int lcpao_grid_set_tables(lcpao_ctx *ctx, const double *const *phi);

/* Uploads the gradient tables dOrbs_Grid[1..3].data[atom]
 * (Set_dOrbitals_Grid.jl:1-20), same layout as phi. */
int lcpao_grid_set_dtables(lcpao_ctx *ctx, const double *const *dx, const double *const *dy,
                           const double *const *dz);

/* Computes the tables on the device: Set_Orbitals_Grid!
 * (Set_Orbitals_Grid.jl:45-277) and, with with_derivatives != 0, also
 * Set_dOrbitals_Grid! (Set_dOrbitals_Grid.jl:23-362).  Needs the lattice
 * fields of lcpao_geom. */
int lcpao_grid_build_tables(lcpao_ctx *ctx, int32_t nspecies, const lcpao_pao *paos, int32_t with_derivatives);

/* Replaces the atom positions (natom x 3, bohr) used by the next
 * lcpao_grid_build_tables; the grid and overlap lists are kept.  No
 * counterpart in the reference (which rebuilds UCell per geometry): it
 * serves finite-difference checks of Force3 with fixed lists. */
int lcpao_grid_set_positions(lcpao_ctx *ctx, const double *gxyz);

/* Copies the device tables back (NO_A x GridN_A per atom; which: 0 = phi,
 * 1..3 = d/dx, d/dy, d/dz). */
int lcpao_grid_get_tables(lcpao_ctx *ctx, int32_t which, double *const *out);

/* Grid Hamiltonian, Calc_MatrixElements_dVH_Vxc_{off,on,nc}!
 * (Hamiltonian.jl:235-380) with _Calc_Hamiltonian_Blocked! (:384-429):
 * nch = 1, 2 or 4 channels V[c] (NN doubles each); H[c] receives the
 * rank-local MPI_Hks[c] (MPI_Hsize doubles), overwritten and already scaled
 * by GridVol. */
int lcpao_grid_hamiltonian(lcpao_ctx *ctx, int32_t nch, const double *const *V, double *const *H,
                           int32_t precision);

/* Density, Set_Density_Grid_{nonpol,pol,nc}! (Set_Density_Grid.jl:60-240):
 * nspin = 1, 2 or 4 global flat density matrices DM[s]; the rank slice
 * starts at dm_offset (MPHks[rank]).  rho[s] (NN doubles) is overwritten
 * with the rank-local partial sum (the caller still reduces over ranks). */
int lcpao_grid_density(lcpao_ctx *ctx, int32_t nspin, const double *const *DM, int64_t dm_offset,
                       double *const *rho, int32_t precision);

/* Force3_{nospin,spin,nc} (Force.jl:407-566): for each channel c, the
 * per-pair sums of _Calc_Force3_Blocked! (Force.jl:3321-3392) with V[c] and
 * DM[c], weighted by weight[c] and GridVol, are ADDED into force
 * (natom x 3, column-major like the Julia Matrix VpotForce).  Reference
 * weights: nospin {4}; spin {2,2}; nc {2,2,4,-4}.  Needs the gradient
 * tables.  LCPAO_BF16X3 runs the block form on the matrix cores (first call:
 * FP32 gradient tiles of 12 bytes per block row and point are built);
 * LCPAO_FP32 the pair-list kernels on FP32 table copies. */
int lcpao_grid_force3(lcpao_ctx *ctx, int32_t nch, const double *const *V, const double *const *DM,
                      int64_t dm_offset, const double *weight, double *force, int32_t precision);

/* Device-resident variants of the three contractions: same semantics, all
 * array pointers are device pointers; d_force stays natom x 3 column-major
 * and is accumulated into.  These are the entry points timed by the
 * benchmark (inputs already in HBM).  They launch on the context's stream
 * (its own non-blocking stream, or the one bound by lcpao_grid_set_stream)
 * and return when it has finished; inputs written by other streams must be
 * complete before the call (synchronise their producer, or bind it). */
int lcpao_grid_hamiltonian_dev(lcpao_ctx *ctx, int32_t nch, const double *const *d_V, double *const *d_H,
                               int32_t precision);
int lcpao_grid_density_dev(lcpao_ctx *ctx, int32_t nspin, const double *const *d_DM, int64_t dm_offset,
                           double *const *d_rho, int32_t precision);
int lcpao_grid_force3_dev(lcpao_ctx *ctx, int32_t nch, const double *const *d_V, const double *const *d_DM,
                          int64_t dm_offset, const double *weight, double *d_force, int32_t precision);

/* Allocation-free FP64 Force3: lcpao_grid_force3_dev with precision
 * LCPAO_FP64 (same arguments, same kernels, same result), except that its
 * per-unit partial sums and per-atom sums live in a fixed 128 MiB array of
 * device memory that the library holds on each GPU for the whole process
 * instead of in context buffers, and that it allocates nothing: no
 * hipMalloc / hipFree, no event timing, no autotuning (a context whose FP64
 * Force3 variant is still open runs the vector-FMA kernel; fix one with
 * lcpao_grid_tune).  No counterpart in the reference: it is for callers that
 * cannot allocate device memory during the call (code under HIP graph
 * capture on the caller's side, tight serving loops, harnesses that forbid
 * allocation).  Limit: 3 doubles per Force3 unit (one unit per pair and 256
 * overlap points) plus 3 per atom must fit in 2^24 doubles, else
 * LCPAO_ERANGE (Si511 at 192^3 needs about 9.6 M).  Every call on one GPU
 * shares that array, so calls that overlap on one GPU (other threads, other
 * streams) race; calls in sequence are safe (the per-atom sums are cleared by
 * a kernel at the start of every call). */
int lcpao_grid_force3_dev_noalloc(lcpao_ctx *ctx, int32_t nch, const double *const *d_V, const double *const *d_DM,
                                  int64_t dm_offset, const double *weight, double *d_force);

/* Makes the context launch on the caller's HIP stream (a hipStream_t passed
 * as void*; NULL restores the context's own stream).  No counterpart in the
 * reference (which has no GPU): it lets a caller order the contractions
 * with its own device work. */
int lcpao_grid_set_stream(lcpao_ctx *ctx, void *hip_stream);

/* Per-kernel timing with HIP events recorded on the launch stream around
 * every kernel of the contractions (no counterpart in the reference; its
 * TimerOutputs labels are Set_Hamiltonian / Set_Density_Grid / Force3).
 * enable != 0 starts recording (and clears the counters); kernel_stats
 * returns the launch count and total milliseconds of one kernel id:
 * 0 H units, 1 H reduce, 2 density fold, 3 density units, 4 density gather,
 * 5 Force3 units, 6 Force3 reduce, 7 atomic densities, 8 FFT passes,
 * 9 Poisson grid kernels (charge, 4 pi / G^2 factor, potential),
 * 10 Bloch sums, 11 density matrix from states. */
int lcpao_grid_profile(lcpao_ctx *ctx, int32_t enable);
int lcpao_grid_kernel_stats(const lcpao_ctx *ctx, int32_t kernel_id, int64_t *launches, double *total_ms);

/* Autotuner (no counterpart in the reference).  Some contractions have
 * several kernels or work-unit shapes that give the same result within the
 * form's tolerance; by default (variant -1) the first call on a context
 * times every candidate once on the context's own geometry and keeps the
 * fastest (a candidate that cannot serve the context is not run and reads
 * -1 ms).  op:
 *   LCPAO_TUNE_RHO64  FP64 density  (0 matrix-core kernel; 1..5 vector-FMA
 *                     kernel with units of (waves per workgroup x grid-list
 *                     entries) 4 x 2048, 4 x 1024, 8 x 512, 2 x 1024, 4 x 512;
 *                     6 block form: vector FMAs over 4x4x4 sub-blocks with
 *                     the orbital columns staged in LDS, 7 the same reading
 *                     the columns from the tables; 8 block form on the
 *                     FP64 matrix cores: per sub-block and slot y, the W
 *                     blocks of y's pairs stacked row-wise times the staged
 *                     orbital columns of y; 6, 7 and 8 are used only when
 *                     the block plan holds the geometry)
 *   LCPAO_TUNE_F364   FP64 Force3   (0 matrix-core kernel, 1 vector-FMA kernel)
 *   LCPAO_TUNE_RHOBF  bf16x3 density (0 single-buffered chunk planes,
 *                     1 double-buffered, used only when its LDS fits;
 *                     2 double-buffered with the chunks loaded straight
 *                     into LDS, used only for blocks of at most 512 rows;
 *                     3 items of up to 8 super-tiles, two workgroups per
 *                     CU, for the blocks of at most 384 rows (the larger
 *                     blocks as in 0); 0-2 form each
 *                     point's sum in the same order, 3 adds the FP32 sums
 *                     of a block's items in FP64)
 *   LCPAO_TUNE_RHO32  FP32 density  (0 pair-list kernel 4 x 2048, 1 block
 *                     form on the FP32 matrix cores, used only when the block
 *                     plan holds the geometry, 2 pair lists 4 x 1024, 3 pair
 *                     lists 8 x 512)
 * lcpao_grid_tune fixes a variant (or -1 to re-tune at the next call);
 * lcpao_grid_tuned reports the choice and, when candidate_ms is not NULL,
 * the measured milliseconds of up to ncand candidates. */
enum { LCPAO_TUNE_RHO64 = 0, LCPAO_TUNE_F364 = 1, LCPAO_TUNE_RHOBF = 2, LCPAO_TUNE_RHO32 = 3 };
int lcpao_grid_tune(lcpao_ctx *ctx, int32_t op, int32_t variant);
int lcpao_grid_tuned(const lcpao_ctx *ctx, int32_t op, int32_t *variant, float *candidate_ms, int32_t ncand);

/* Autotuner cache (no counterpart in the reference): a JSON file of chosen
 * variants keyed by the context's shape (its distinct orbital counts, the
 * grid spacing GridVol^(1/3) and the mean pairs per atom), so that a later
 * run on a like geometry skips the timing.  lcpao_grid_tune_load fixes every
 * operation stored under this context's key (*loaded = how many; a missing
 * file loads nothing and is not an error); lcpao_grid_tune_save writes the
 * context's tuned operations into the file, replacing its earlier entries
 * for the same key and keeping the others. */
int lcpao_grid_tune_load(lcpao_ctx *ctx, const char *path, int32_t *loaded);
int lcpao_grid_tune_save(lcpao_ctx *ctx, const char *path);

/* ---- grid fields next to the contractions --------------------------------
 * Radial data of one species for the atomic densities: the PAO log mesh and
 * valence density (Spe_PAO_XV/RV, Spe_Atomic_Den of Read_PAO.jl:183-303, the
 * density already renormalised to the species' core charge) and the VPS log
 * mesh and partial-core density (Spe_VPS_XV/RV, Spe_Atomic_PCC of
 * Read_VPS.jl:196-272).  The two density arrays hold nmesh + 4 values, mesh
 * point k at index k + 1, with the reference's linear end points. */
typedef struct {
    int32_t nmesh_pao;
    const double *pao_xv, *pao_rv; /* [nmesh_pao] */
    const double *atomic_den;      /* [nmesh_pao + 4] */
    int32_t nmesh_vps;
    const double *vps_xv, *vps_rv; /* [nmesh_vps] */
    const double *atomic_pcc;      /* [nmesh_vps + 4] */
} lcpao_atomic_radial;

/* Set_AdenPCC_Grid (Set_AdenPCC_Grid.jl:1-142): atomic density ADensity,
 * partial-core density PCCDensity (NN each) and the initial spin densities
 * rho[nspin] (NN each; nspin 1 "off", 2 "on", 4 "nc", the nc rotation of
 * :125-138 applied) from the atoms' grid lists.  init_nspin: natom x 2 (up,
 * down electrons, Init_Atoms_Nspin); init_angle: natom x 2 (theta, phi,
 * Init_Atoms_Angle[1:2]; only read for nspin 4, may be NULL otherwise).
 * Needs the lattice, positions, species and cell lists of lcpao_geom.  Every
 * atom of the context contributes, so the result is the reference's after
 * its MPI.Allreduce! (:118-122).  Each grid point's sum runs in the
 * reference's atom order.  The _dev variant takes device output pointers. */
int lcpao_grid_aden_pcc(lcpao_ctx *ctx, int32_t nspin, int32_t nspecies, const lcpao_atomic_radial *spe,
                        const double *init_nspin, const double *init_angle, double *aden, double *pcc,
                        double *const *rho);
int lcpao_grid_aden_pcc_dev(lcpao_ctx *ctx, int32_t nspin, int32_t nspecies, const lcpao_atomic_radial *spe,
                            const double *init_nspin, const double *init_angle, double *d_aden, double *d_pcc,
                            double *const *d_rho);

/* Solve_Poisson! (Poisson.jl:1-108): dVHart (NN) of the charge
 * Density_xyz = 2 rho[0] - 2 aden (nspin 1) or rho[0] + rho[1] - 2 aden
 * (nspin 2 or 4), as Set_Density_xyz! forms it; nspin 0 solves for rho[0]
 * itself (the inner Solve_Poisson!(Ngrid, Recvecs, ...), aden unused).
 * V(G) = 4 pi n(G) / |G|^2 with the reference's signed indices (index i of
 * an axis of N points counts as i - N from N / 2 on, integer division, so
 * for odd N the middle index is negative), V(0) = 0, real part of the
 * normalised inverse transform.  recvecs: Recvecs = 2 pi inv(Latvecs') in
 * Julia's column-major order (recvecs[3 j + i] = Recvecs[i+1, j+1]).  FFTs
 * by the library's own LDS-resident mixed-radix passes (axes of at most 1024
 * points, prime factors up to 13).  The _dev variant takes device
 * pointers for rho, aden and vhart. */
int lcpao_grid_poisson(lcpao_ctx *ctx, int32_t nspin, const double *const *rho, const double *aden,
                       const double *recvecs, double *vhart);
int lcpao_grid_poisson_dev(lcpao_ctx *ctx, int32_t nspin, const double *const *d_rho, const double *d_aden,
                           const double *recvecs, double *d_vhart);

/* ---- either side of the k-point diagonalisation -------------------------
 * Dense matrices are fsize x fsize complex (fsize = sum of Total_NumOrbs),
 * column-major with interleaved real / imaginary parts, as Julia's
 * Matrix{ComplexF64}; row MP[A] + i is orbital i of atom A.  Both routines
 * cover the pairs held by the context (all pairs when it was built for one
 * rank); flat arrays are the context's my_hsize pair blocks.
 *
 * HS_matrix! with a k-point (HS_matrix.jl:88-110, 143-166):
 * dense[m][A_i, B_j] = sum over the pairs (A, B, cell) of
 * flat[m][pair][i, j] * cispi(2 k . atv_ijk[cell]); kpt = (ka, kb, kc) in
 * the reciprocal basis; nmat 1..4 matrices (e.g. S and H) in one pass.
 * Each element's images are summed in the pair-list order. */
int lcpao_grid_bloch_sum(lcpao_ctx *ctx, const double *kpt, int32_t nmat, const double *const *flat,
                         double *const *dense);
int lcpao_grid_bloch_sum_dev(lcpao_ctx *ctx, const double *kpt, int32_t nmat, const double *const *d_flat,
                             double *const *d_dense);
/* fsize and the number of distinct (A, B) atom pairs of the context. */
int lcpao_grid_bloch_sizes(lcpao_ctx *ctx, int64_t *fsize, int64_t *ngroups);

/* Calc_DM_Crystal_Collinear_nopol! (Crystal_DFT.jl:219-270) from the states:
 * DM[pair][i, j] = sum over the nk k-points (in order) of
 * Re(DM_k[B_j, A_i] * cispi(2 k . atv_ijk[cell]) * scale), with
 * DM_k = C[k] C[k]^H formed only for the (A, B) blocks the pairs need
 * (no fsize x fsize matrix).  C[k]: fsize x nstates complex, column-major,
 * columns scaled by sqrt(occupation * k weight) as Calc_fnkCnk! leaves
 * them; kpts: nk x 3; scale: 1 / AllNkpt.  DM (my_hsize) is overwritten;
 * the reference's MPI.Allreduce! over k-point ranks stays the caller's. */
int lcpao_grid_dm_states(lcpao_ctx *ctx, int32_t nk, const double *kpts, int64_t nstates, const double *const *C,
                         double scale, double *DM);
int lcpao_grid_dm_states_dev(lcpao_ctx *ctx, int32_t nk, const double *kpts, int64_t nstates,
                             const double *const *d_C, double scale, double *d_DM);
/* Calc_DM_Crystal_Collinear_pol! (Crystal_DFT.jl:276-339) is two calls of
 * lcpao_grid_dm_states, one per spin (Cnk[1] -> DM[1], Cnk[2] -> DM[2]). */

/* HS_matrix_NC! with a k-point, flat Hks (HS_matrix.jl:211-257): H is one
 * (2 fsize) x (2 fsize) complex matrix, column-major, overwritten:
 *   H[A_i, B_j]                 = sum (hks[0] + im ihks[0]) * ex   (up-up)
 *   H[fsize+A_i, fsize+B_j]     = sum (hks[1] + im ihks[1]) * ex   (down-down)
 *   H[A_i, fsize+B_j]           = sum (hks[2] + im (hks[3] + ihks[2])) * ex
 * with ex = cispi(2 k . atv_ijk[cell]); the down-up block stays zero as in
 * the reference.  hks: the four flat Hks channels (my_hsize each); ihks: the
 * three nested iHks arrays flattened in pair order (atom, Rn, ist, jst),
 * which is the flat pair-block layout. */
int lcpao_grid_bloch_sum_nc(lcpao_ctx *ctx, const double *kpt, const double *const *hks, const double *const *ihks,
                            double *H);
int lcpao_grid_bloch_sum_nc_dev(lcpao_ctx *ctx, const double *kpt, const double *const *d_hks,
                                const double *const *d_ihks, double *d_H);

/* Calc_DM_Crystal_NonCollinear! (Crystal_DFT.jl:342-410) from the spinor
 * states: C[k] is (2 fsize) x nstates complex, column-major, scaled as
 * Calc_fnkCnk! leaves it.  With P = C C^H formed per (A, B) block only:
 *   DM[0] = sum_k Re(P[B_j, A_i] ex)              (up-up)
 *   DM[1] = sum_k Re(P[fsize+B_j, fsize+A_i] ex)  (down-down)
 *   DM[2] = sum_k Re(P[fsize+B_j, A_i] ex), DM[3] = sum_k Im(P[fsize+B_j, A_i] ex)
 * with ex = cispi(2 k . atv_ijk[cell]) * scale.  iDM (two pointers, or NULL)
 * receives the k-point sums of Calc_iDM_Crystal_NonCollinear!
 * (Crystal_DFT.jl:413-486): iDM[0] = sum_k Im(P[B_j, A_i] ex), iDM[1] the
 * same on the down-down block, in the flat pair order of its nested arrays.
 * All outputs (my_hsize each) are overwritten; the MPI.Allreduce! stays the
 * caller's. */
int lcpao_grid_dm_states_nc(lcpao_ctx *ctx, int32_t nk, const double *kpts, int64_t nstates, const double *const *C,
                            double scale, double *const *DM, double *const *iDM);
int lcpao_grid_dm_states_nc_dev(lcpao_ctx *ctx, int32_t nk, const double *kpts, int64_t nstates,
                                const double *const *d_C, double scale, double *const *d_DM, double *const *d_iDM);

/* Which kernels serve the split-precision forms (no counterpart in the
 * reference).  The block forms (bf16x3 H, density and Force3 on the matrix
 * cores, the FP32 H and the FP32 block density) work on 8x8x8-point grid
 * blocks whose rows are the orbitals of every sphere image covering the
 * block; they hold at most 32 orbitals per atom, 576 rows and 45 sphere
 * images per block.  A geometry beyond that is served by the pair-list
 * kernels of a wider form, which meet the narrower form's bound: the FP32
 * density and Force3 for bf16x3, the FP64 H for FP32 and bf16x3.
 * *state: -1 not decided yet (no split-precision call so far), 1 block
 * forms, 0 pair-list kernels. */
int lcpao_grid_block_forms(const lcpao_ctx *ctx, int32_t *state);

/* Host-only check of the block-form limits above for a geometry (no device
 * needed): the largest padded row count and sphere-image count over the
 * 8x8x8 blocks and the number of blocks.  Returns LCPAO_OK when the block
 * forms hold the geometry, LCPAO_ERANGE when the pair-list kernels will
 * serve it, LCPAO_EINVAL for a geometry without cell lists. */
int lcpao_grid_block_layout(const lcpao_geom *geom, int32_t *max_rows, int32_t *max_slots, int64_t *nblocks);

/* Sizes of the context (rank-local H size, grid size, sums of the lists). */
int lcpao_grid_sizes(const lcpao_ctx *ctx, int64_t *my_hsize, int64_t *nn, int64_t *total_olg,
                     int64_t *total_gridn);

/* Message of the last failure on this context (or of the last failed
 * lcpao_grid_create when ctx is NULL). */
const char *lcpao_grid_last_error(const lcpao_ctx *ctx);

/* Frees the context and its device memory. */
void lcpao_grid_destroy(lcpao_ctx *ctx);

/* ------------------------------------------------------------------ */
/* Host-side UCell list construction (UCell.jl:210-294): the neighbour  */
/* search Get_FNAN/Trn_System (:520-672), Calc_AtomsGrid (:297-411),    */
/* split_system_grid (:76-132) and Calc_AtomOverlap_Grid (:425-516).    */
/* ------------------------------------------------------------------ */
typedef struct lcpao_ucell lcpao_ucell;

/* ---- batched form (no counterpart in the reference) ------------------------
 * The split-BF16x3 (or FP32) grid H (1 channel) and density (1 spin) of many
 * structures in one go: the block items of all structures are merged into
 * one list per operation, run as one launch (persistent, its workgroups
 * pulling items from a work queue, or one workgroup per item); the whole
 * sequence can be captured in a HIP graph and replayed.  For callers that assemble many
 * small or uneven cells (structure screening, many geometries).
 *
 * lcpao_batch_create takes n contexts on one device, each with tables set
 * and the whole pair list of its structure (one rank); it builds their block
 * plans.  The contexts must outlive the batch and must not be used from
 * another thread while it runs.
 * lcpao_batch_assemble: per structure s, d_V[s] (NN_s), d_DM[s] (flat pair
 * blocks, Total_Hsize_s), outputs d_H[s] (scaled by GridVol) and d_rho[s]
 * (overwritten); all device pointers.  use_graph != 0 replays a graph
 * captured at the first call with these pointers (re-captured when they
 * change).  The first call also times the schedules (0 persistent work
 * queue, 1 one workgroup per item, 2 as 1 with the density of the blocks
 * of at most 384 rows in items of at most 8 super-tiles, two workgroups per
 * CU) and keeps the fastest (lcpao_batch_tune fixes one, -1 re-tunes;
 * lcpao_batch_tuned writes the three candidates' milliseconds).  Returns
 * after the stream has been synchronised. */
typedef struct lcpao_batch lcpao_batch;
int lcpao_batch_create(lcpao_batch **out, int32_t n, lcpao_ctx *const *ctxs);
int lcpao_batch_assemble(lcpao_batch *b, const double *const *d_V, const double *const *d_DM, double *const *d_H,
                         double *const *d_rho, int32_t use_graph);
/* Precision of the batched assembly: LCPAO_BF16X3 (default) or LCPAO_FP32
 * (the FP32 block kernels: H on the FP32 matrix cores, the FP32 block
 * density; one workgroup per item).  A change drops the captured graph; the
 * next lcpao_batch_assemble builds the FP32 tiles of every member. */
int lcpao_batch_set_precision(lcpao_batch *b, int32_t precision);
int lcpao_batch_tune(lcpao_batch *b, int32_t sched);
int lcpao_batch_tuned(const lcpao_batch *b, int32_t *sched, float *candidate_ms);
int lcpao_batch_sizes(const lcpao_batch *b, int32_t *nstruct, int32_t *nclasses, int64_t *nitems);
const char *lcpao_batch_last_error(const lcpao_batch *b);
void lcpao_batch_destroy(lcpao_batch *b);

/* latvecs: 3x3 row-major (rows are lattice vectors); gxyz natom x 3;
 * cut[natom] = Atoms_Cut1; no[natom] = Total_NumOrbs (may be NULL);
 * nprocs/rank select the rank-local pair range. */
int lcpao_ucell_build(lcpao_ucell **out, const double *latvecs, int32_t natom, const double *gxyz,
                      const double *cut, const int32_t *no, const int32_t *ngrid, const double *origin,
                      int32_t nprocs, int32_t rank, int32_t nthreads);

/* Scalars of the built UCell. */
int lcpao_ucell_scalars(const lcpao_ucell *u, int32_t *cpycell, int64_t *nloop, int64_t *npairs,
                        int64_t *total_hsize, int64_t *my_hsize, int64_t *my_hoff, double *gridvol,
                        int64_t *total_gridn, int64_t *total_olg);

/* Copies the lists into caller arrays (sizes from lcpao_ucell_scalars; any
 * pointer may be NULL to skip that list).  Neighbour lists are CSR by atom
 * (nb_off[natom+1]); grid lists CSR by atom (gl_off[natom+1]); pair arrays
 * have npairs entries with 0-based atoms; olg lists CSR by pair
 * (olg_off[npairs+1]). */
int lcpao_ucell_copy(const lcpao_ucell *u, int32_t *fnan, int64_t *nb_off, int32_t *natn, int32_t *ncn,
                     int32_t *gridn, int64_t *gl_off, int32_t *glist, int32_t *clist, int32_t *pair_atom,
                     int32_t *pair_natn, int32_t *pair_ncn, int64_t *pair_hoff, int32_t *numolg,
                     int64_t *olg_off, int32_t *olg1, int32_t *olg2);

const char *lcpao_ucell_last_error(void);
void lcpao_ucell_destroy(lcpao_ucell *u);

#ifdef __cplusplus
}
#endif
#endif /* LCPAO_GRID_H */
