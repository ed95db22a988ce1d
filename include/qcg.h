/*
 * qcg.h — C ABI of the MI355X (gfx950) Fock-build engine.
 *
 * Each entry point replaces one function of ExaPsi/IQCP qc-core's hot path (paths relative to
 * crates/qc-core/src/).  Plain C types only: caller-owned host buffers (or device pointers for
 * the *_device variants), every call returns a status code (0 = OK), qcg_last_error() gives the
 * message.  One context per host thread; device memory is owned by the context.
 *
 * Matrix layout: nbf x nbf, row-major.  D, J, K, F, S, H are symmetric, so row- and
 * column-major (qc-core's nalgebra column-major ScfOutput fields, scf/mod.rs:1567-1569) agree.
 * Basis-function order: shells in the order given to qcg_set_basis, Cartesian components
 * xx, xy, xz, yy, yz, zz (integrals/cartesian.rs:113-141).  Coordinates in bohr.
 */
#ifndef QCG_H
#define QCG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (ScfError / RysError / BoysError mapping, scf/mod.rs:73-109, rys/mod.rs:82-114). */
#define QCG_OK 0
#define QCG_INVALID_ARG 1    /* DimensionMismatch / InvalidBasisSize / InvalidOrder / NegativeArgument */
#define QCG_UNSUPPORTED_L 2  /* AngularMomentumTooHigh (l > 2) */
#define QCG_NUMERICAL 3      /* NumericalInstability / SingularOverlap */
#define QCG_DEVICE 4         /* HIP runtime error */
#define QCG_NOT_CONVERGED 5  /* ScfError::NotConverged */
#define QCG_NO_BASIS 6
#define QCG_NO_GRID 7

/* Precision of the two-electron kernels. */
#define QCG_PREC_FP64 0   /* every surviving quartet in FP64 */
#define QCG_PREC_MIXED 1  /* quartets with Q_ab*Q_cd < fp32_cutoff in FP32, the rest FP64; FP64 accumulation */

typedef struct qcg_context qcg_context;

/* Convergence/acceleration settings: ScfConfig + ConvergenceProfile (scf/mod.rs:123-262). */
typedef struct qcg_scf_config {
    double energy_tol;   /* Tight 1e-10, Medium 1e-8, Loose 1e-6 */
    double density_tol;  /* Tight 1e-8,  Medium 1e-6, Loose 1e-4 */
    int max_iterations;  /* default 100 */
    int use_diis;        /* AO-basis DIIS */
    int diis_size;       /* default 8 */
    int diis_start;      /* default 1 */
    double damp;         /* default 0 */
    int damp_start;      /* default 5 */
    double level_shift;  /* default 0 */
} qcg_scf_config;

/* Scalar part of ScfOutput (scf/mod.rs:629-650). */
typedef struct qcg_scf_result {
    int converged;
    int iterations;
    double energy_total;
    double energy_electronic;
    double energy_nuclear;
    double last_delta_e;
    double last_rms_d;
    double energy_xc;    /* KS only */
    double energy_j;     /* KS only: 1/2 Tr DJ */
    double energy_k;     /* KS only: -1/4 a Tr DK */
    double energy_1e;    /* KS only: Tr DH */
} qcg_scf_result;

/* One SCF iteration, ScfIteration (scf/mod.rs:627-650): iteration 0 is the guess density's build
 * (delta_e and rms_density_change NaN there, the reference's None); diis_applied: the density of
 * this iteration came from a DIIS-extrapolated Fock matrix.  Two fields have no counterpart in
 * qc-core, which builds in FP64 only: fp32_build = 1 when this iteration's Fock build ran mixed
 * precision, and error_bound = qcg_error_bound's bound on that build's two-electron energy error
 * (0 for FP64 builds). */
typedef struct qcg_scf_iteration {
    int iteration;
    int converged;
    int diis_applied;
    int fp32_build;
    double energy_total;
    double energy_electronic;
    double delta_e;
    double rms_density_change;
    double error_bound;
} qcg_scf_iteration;

/* ---- context ---------------------------------------------------------------------------- */
int qcg_create(qcg_context** ctx, int device);
void qcg_destroy(qcg_context* ctx);
const char* qcg_last_error(const qcg_context* ctx);

/* Basis set: replaces BasisSet::build (basis/mod.rs:164) output.  Shell s has angular momentum
 * l[s] (0..2), primitives [prim_off[s], prim_off[s]+nprim[s]) of exps/coefs (raw contraction
 * coefficients, basis/primitives.rs:232-237), centre centers[3s..3s+2], atom shell_atom[s].
 * Z/xyz: nuclei (for V, E_nuc, grids, gradients).  Builds shell pairs, Schwarz bounds on the
 * device (compute_schwarz_bounds, integrals/eri/mod.rs:1263) and the class-sorted pair lists. */
int qcg_set_basis(qcg_context* ctx, int nshell, const int* l, const int* nprim, const int* prim_off,
                  const double* exps, const double* coefs, const double* centers, const int* shell_atom,
                  int natom, const int* Z, const double* xyz);

/* A batch of independent molecules (replaces nmol BasisSet::build calls): molecule m owns the next
 * mol_nshell[m] shells and mol_natom[m] atoms of the concatenated arrays; shell_atom is the atom
 * index inside the molecule.  Quartets never mix molecules.  Matrices of the batch entry points are
 * packed block-diagonally: molecule m's nbf_m x nbf_m block follows those of molecules < m. */
int qcg_set_basis_batch(qcg_context* ctx, int nmol, const int* mol_nshell, const int* mol_natom, const int* l,
                        const int* nprim, const int* prim_off, const double* exps, const double* coefs,
                        const double* centers, const int* shell_atom, const int* Z, const double* xyz);

/* precision: QCG_PREC_*; fp32_cutoff: Schwarz-product cutoff below which quartets run in FP32
 * (default 1e-6); schwarz_threshold: SCHWARZ_THRESHOLD (integrals/eri/mod.rs:1258, default 1e-12).
 * QCG_PREC_MIXED also runs the XC grid products in FP32 (qcg_set_xc_mode); the SCF drivers
 * (qcg_rhf_scf, qcg_ks_scf) switch to FP64 builds once the density change drops below 1e-5 and
 * test convergence on FP64 builds only, so converged energies are those of the FP64 operator. */
int qcg_set_options(qcg_context* ctx, int precision, double fp32_cutoff, double schwarz_threshold);
int qcg_nbf(const qcg_context* ctx, int* nbf);
/* Molecular charge (electron count = sum Z - charge), PresetSystem.nelec (scf/mod.rs:284-310). */
int qcg_set_charge(qcg_context* ctx, int charge);

/* ---- special functions (batched, one argument per lane) -------------------------------- */
/* boys_eval_all (boys/mod.rs:327): F[i*(m_max+1)+m] = F_m(T[i]). */
int qcg_boys_eval_all(qcg_context* ctx, int m_max, int n, const double* T, double* F);
/* rys_roots (rys/mod.rs:253): roots u = t^2 in [0,1) and weights, [i*nroots+k]. nroots 1..7. */
int qcg_rys_roots(qcg_context* ctx, int nroots, int n, const double* T, double* roots, double* weights);

/* ---- two-electron integrals and J/K ----------------------------------------------------- */
/* compute_schwarz_bounds (integrals/eri/mod.rs:1263): Q[i*nshell+j] = sqrt(max |(ij|ij)|). */
int qcg_schwarz_bounds(qcg_context* ctx, double* Q);
/* Integer screening outputs: unique shell quartets (8-fold) with Q_ij*Q_kl >= threshold, and how
 * many of those run in FP32 under QCG_PREC_MIXED. */
int qcg_screen_counts(qcg_context* ctx, int64_t* n_kept, int64_t* n_fp32);
/* No counterpart in qc-core (it computes every quartet in FP64): an a-priori bound on the energy
 * error of a mixed-precision J/K build at density D against the all-FP64 build of the same
 * quartets,
 *   |E2(mixed) - E2(FP64)| <= eps32 * sum over FP32 quartets (ab|cd) of
 *                             Q_ab Q_cd deg [4 S_ab S_cd + hf_frac (S_ac S_bd + S_ad S_bc)],
 * E2 = 1/2 Tr D (J - hf_frac/2 K), Q the Schwarz bounds (|(ij|kl)| <= Q_ab Q_cd per component),
 * S_xy the sum of |D| over the shell block (x, y), deg the 8-fold symmetry weight of the quartet,
 * and eps32 = 2^-16 the relative error allowed per FP32 quartet (a single quartet measures below
 * 2^-20).  *dE_bound = 0 in FP64 mode.  Single molecules only (QCG_INVALID_ARG after
 * qcg_set_basis_batch). */
int qcg_error_bound(qcg_context* ctx, const double* D, double hf_frac, double* dE_bound);
/* The per-iteration trace of the context's last qcg_rhf_scf / qcg_ks_scf (ScfOutput.trace,
 * scf/mod.rs:652-700): copies min(cap, length) entries to out, *n = the trace length. */
int qcg_scf_history(const qcg_context* ctx, qcg_scf_iteration* out, int cap, int* n);
/* eri_compressed (integrals/eri/mod.rs:992): out has npair(npair+1)/2 doubles, npair = nbf(nbf+1)/2,
 * index eri_index (scf/mod.rs:608-613). */
int qcg_eri_compressed(qcg_context* ctx, double* out);
/* build_jk (dft/ks_scf.rs:1369) / build_coulomb (:1306) / build_exchange (:1342), integral-direct:
 * J_mn = sum_ls D_ls (mn|ls), K_mn = sum_ls D_ls (ml|ns).  J or K may be NULL. */
int qcg_build_jk(qcg_context* ctx, const double* D, double* J, double* K);
/* build_fock (scf/mod.rs:825): F = H + J - K/2. */
int qcg_build_fock(qcg_context* ctx, const double* H, const double* D, double* F);
/* No counterpart in qc-core: for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same J/K build as
 * qcg_build_jk_device, and bit-identical to it, with its accumulators in a static device array of
 * 4.2 MB (one molecule of at most 512 basis functions; larger ones get QCG_INVALID_ARG) instead of
 * the context's buffers.  It allocates nothing: the context's work plan must exist for the current
 * options and tiles (one qcg_build_jk_device call after the last change), else QCG_INVALID_ARG.
 * Every call on one GPU shares that GPU's array, so calls that overlap on one GPU race. */
int qcg_build_jk_device_noalloc(qcg_context* ctx, const double* dD, double* dJ, double* dK, void* stream);
/* Device-pointer variant of qcg_build_jk on the given HIP stream (NULL = context stream);
 * asynchronous with respect to the host. */
int qcg_build_jk_device(qcg_context* ctx, const double* dD, double* dJ, double* dK, void* stream);
/* Device-pointer Fock build: dF = dH + J - k_factor K (RHF build_fock: k_factor = 1/2; hybrid
 * KS, dft/ks_scf.rs:1960: k_factor = a/2). */
int qcg_build_fock_device(qcg_context* ctx, const double* dH, const double* dD, double k_factor, double* dF,
                          void* stream);
/* Graph replay for qcg_build_fock_device: on != 0 captures the whole launch sequence of a device
 * Fock build (fill, class kernels on the auxiliary streams, combine) as one HIP graph the first
 * time and replays it afterwards; it is captured again when the buffers, stream, k factor,
 * precision options or tiles change.  Profiling mode always launches eagerly.  No counterpart in
 * qc-core: launch mode only. */
int qcg_set_graph(qcg_context* ctx, int on);

/* Per quartet-class statistics of the integral-direct J/K build: integer counts (exact, same
 * screens as the kernels), the algorithmic flop model, and (after qcg_set_profiling(ctx, 1) and
 * one build) the measured kernel time of the last build.  Flop model per surviving primitive
 * quartet of class (la lb|lc ld) with n = (la+lb+lc+ld)/2+1 Rys nodes and N_q Cartesian
 * components:  n [80 + 15 (la+lb+1)(lc+ld+1) + N_q (4 + 6 (lb+ld))], plus 12 N_q per contracted
 * quartet for the J/K digestion.  Classes are ordered bra class 5..0, ket class <= bra. */
typedef struct qcg_class_stat {
    int la, lb, lc, ld;
    int64_t quartets_fp64, quartets_fp32;
    int64_t prim_quartets_fp64, prim_quartets_fp32;
    double flops_fp64, flops_fp32;
    double ms_fp64, ms_fp32;
} qcg_class_stat;
int qcg_set_profiling(qcg_context* ctx, int on);
/* Autotuner for the J/K kernels: for every quartet class and precision part (FP64 quartets, FP32
 * quartets of the mixed mode), time each candidate tile (threads per block x kets per thread, or
 * block 0 = one wavefront per contracted quartet, kets per work item) on the device density dD
 * (iters timed builds per candidate) and keep the fastest.  chosen (optional, 84 ints) receives
 * (block, kets) per class and part: [class][part][2]. */
int qcg_autotune(qcg_context* ctx, const double* dD, int iters, int* chosen);
/* No counterpart (launch configuration).  After qcg_autotune: greedy over the classes, longest
 * first, and over the same candidates, changing one class's FP64-part tile at a time and keeping
 * a change when `builds` whole builds (all classes, four streams) run at least 0.5 % faster in two
 * measurements; in mixed precision it also tries each class's FP32 items in a launch of their own
 * (with the FP32-part tile).  chosen (optional) as for qcg_autotune. */
int qcg_autotune_refine(qcg_context* ctx, const double* dD, int builds, int* chosen);
/* Set / read the J/K tiles directly (84 ints, same layout as qcg_autotune's chosen; classes in the
 * order (dd|dd), (dd|dp), ..., (ss|ss)).  No counterpart in qc-core: launch configuration only.
 * Block 0 on a class with fewer than 27 components runs the per-lane kernel with 64 threads; classes
 * with more than 81 components always run the wave-cooperative kernel (kets per wave = the tile's
 * kets for block 0, block x kets otherwise). */
int qcg_set_tiles(qcg_context* ctx, const int* tiles);
int qcg_get_tiles(qcg_context* ctx, int* tiles);
/* Launch schedule of the J/K class kernels.  A build launches the 21 class kernels longest first,
 * each onto the auxiliary stream with the least cost so far, by a cost per class (21 doubles,
 * classes in the order of qcg_set_tiles): the per-class times qcg_autotune measured, or costs set
 * here (e.g. saved with a set of tuned tiles), or, when neither is set (all zero or cost NULL
 * resets), the work plan's model flops of each class's surviving quartets before the primitive
 * screen.  qcg_get_class_cost returns the costs in use (*measured: 1 if set or measured, 0 for
 * the plan's estimate, which exists after the first J/K build).  No counterpart in qc-core:
 * launch configuration only; J and K do not depend on it (order-independent accumulation). */
int qcg_set_class_cost(qcg_context* ctx, const double* cost);
int qcg_get_class_cost(qcg_context* ctx, double* cost, int* measured);
/* Primitive split of the J/K and gradient work items: a bra pair with more than `split` primitive
 * pairs is processed as ranges of at most `split` of them (the integral is linear in the bra
 * primitives), so heavy contracted quartets do not serialise on one lane.  split = 0 turns it off.
 * No counterpart in qc-core: launch configuration only; results agree to rounding. */
int qcg_set_prim_split(qcg_context* ctx, int split);
/* out must hold 21 entries; *n receives the number written. */
int qcg_class_stats(qcg_context* ctx, qcg_class_stat* out, int* n);

/* ---- exchange-correlation (B3LYP) on a Becke/Lebedev grid ------------------------------ */
/* build_becke_grid (dft/grid.rs:103) for the context's atoms: n_radial radial points (75),
 * quality 0 = Standard (302) / 1 = Fine (590), pruning 1 = SG-1.  Also evaluates Phi and grad Phi
 * on the device (evaluate_basis_and_gradients_on_grid, dft/ks_scf.rs:115).  *npts = grid size. */
int qcg_build_grid(qcg_context* ctx, int n_radial, int quality, int pruning, int64_t* npts);
/* A caller-supplied grid (xyz[3*npts] bohr, weights). */
int qcg_set_grid(qcg_context* ctx, int64_t npts, const double* xyz, const double* w);
/* Copy the current grid out (any pointer may be NULL). */
int qcg_get_grid(qcg_context* ctx, double* xyz, double* w, int* atom_idx);
/* B3LYP pointwise (B3lyp::eval_xc_gga, dft/functionals.rs:1625): exc per electron, vrho, vsigma. */
int qcg_xc_eval(qcg_context* ctx, int64_t n, const double* rho, const double* sigma, double* exc, double* vrho,
                double* vsigma);
/* V_xc (symmetrised) and E_xc for density D (build_vxc_gga_matmul, dft/ks_scf.rs:1150). */
int qcg_xc(qcg_context* ctx, const double* D, double* Vxc, double* Exc);
/* Device-pointer variant; *dExc is written on the device.  It uses none of the J/K entry points'
 * buffers, so it may run on another stream concurrently with qcg_build_jk_device on the same
 * context and density (qcg_ks_scf's Fock build does so); two XC builds of one context must not
 * overlap. */
int qcg_xc_device(qcg_context* ctx, const double* dD, double* dVxc, double* dExc, void* stream);
/* XC build kernels.  The hand-written MFMA kernels form X = Phi D block by block and reduce it into
 * rho on the fly, then V = Phi^T aow by tiles; the library form runs DGEMMs around a point kernel.
 * mode 0 (default): MFMA kernels with FP32 products (FP32 accumulation, rho, the functional and
 * E_xc in FP64, plus the first-order FP64 correction of E_xc for rounding D to FP32) for the mixed
 * precision option, and MFMA kernels with FP64 products for FP64 up to 352 basis functions (the
 * library form above that); mode 1: library form for both options; mode 2: MFMA kernels for both
 * (FP64 products for FP64).  No counterpart in qc-core: kernel selection only. */
int qcg_set_xc_mode(qcg_context* ctx, int mode);
/* ks_scf (dft/ks_scf.rs:1567) for B3LYP on the current grid: F = H + J - (a/2) K + V_xc. */
int qcg_ks_scf(qcg_context* ctx, const qcg_scf_config* cfg, const double* D_init, qcg_scf_result* res, double* D,
               double* C, double* eps, double* F);

/* ---- one-electron terms and SCF ---------------------------------------------------------- */
/* overlap_matrix (integrals/overlap.rs:647), kinetic and nuclear attraction (hcore_matrix,
 * integrals/hcore.rs:103 = T + V), computed on the device.  Any output may be NULL.  One molecule
 * per context (a batch context is QCG_INVALID_ARG). */
int qcg_one_electron(qcg_context* ctx, double* S, double* T, double* V);
/* rhf_scf (scf/mod.rs:1363) with the integral-direct GPU Fock build.  D_init (optional) seeds the
 * density (rhf_scf_with_guess, :1610).  Outputs (optional): D, C (columns = MOs), eps, F. */
int qcg_rhf_scf(qcg_context* ctx, const qcg_scf_config* cfg, const double* D_init, qcg_scf_result* res,
                double* D, double* C, double* eps, double* F);

/* ---- preset systems: a stored compressed tensor instead of a basis ---------------------- */
/* build_fock (scf/mod.rs:825-881) as the reference calls it: F = H + sum_ls D_ls [(mn|ls) -
 * 1/2 (ml|ns)] from the compressed 8-fold tensor eri (eri_len = npair(npair+1)/2, npair =
 * nbf(nbf+1)/2, eri_index order, scf/mod.rs:608).  H, D, F: nbf x nbf, D symmetric.  The context
 * needs no basis; a wrong eri_len is QCG_INVALID_ARG ("Invalid ERI array size"). */
int qcg_build_fock_stored(qcg_context* ctx, int nbf, const double* H, const double* D, const double* eri,
                          int64_t eri_len, double* F);
/* rhf_scf (scf/mod.rs:1363) on a PresetSystem (:284): S, H row-major nbf x nbf, the compressed
 * tensor, nelec, e_nuc; validated as PresetSystem::validate (:359).  Outputs as qcg_rhf_scf. */
int qcg_rhf_scf_stored(qcg_context* ctx, int nbf, int nelec, double e_nuc, const double* S, const double* H,
                       const double* eri, int64_t eri_len, const qcg_scf_config* cfg, const double* D_init,
                       qcg_scf_result* res, double* D, double* C, double* eps, double* F);

/* ---- batches of molecules (after qcg_set_basis_batch) ------------------------------------ */
/* J, K for every molecule of the batch in one set of launches; D, J, K packed block-diagonally. */
int qcg_batch_build_jk(qcg_context* ctx, const double* D, double* J, double* K);
/* rhf_scf for every molecule, iterating together: one batched J/K build per iteration (a captured
 * HIP graph replayed each iteration when use_graph != 0).  res[nmol]; D (optional, packed);
 * *ms_jk (optional) = total device time of the J/K builds. */
int qcg_batch_rhf_scf(qcg_context* ctx, const qcg_scf_config* cfg, int use_graph, qcg_scf_result* res, double* D,
                      double* ms_jk);

/* ---- analytic nuclear gradient ---------------------------------------------------------- */
/* Two-electron part (eri_gradient_fused, scf/gradient.rs:1430): grad[3*natom] += sum over unique
 * Schwarz-surviving quartets of [1/2 D_pq D_rs - (hf_frac/4) D_pr D_qs] d(pq|rs)/dR.  grad is
 * overwritten (not accumulated). */
int qcg_eri_gradient(qcg_context* ctx, const double* D, double hf_frac, double* grad);
/* One-electron + nuclear part: dVnn/dR + Tr(D dH/dR) - Tr(W dS/dR) (rhf_gradient :915-950), the
 * one-electron terms on the device.  grad overwritten. */
int qcg_one_electron_gradient(qcg_context* ctx, const double* D, const double* W, double* grad);
/* rhf_gradient (scf/gradient.rs:900): W = 2 sum_i<n_occ eps_i C_i C_i^T (:101), then both parts.
 * C: nbf x nbf, columns = MOs (row-major storage, C[mu*nbf + i]). */
int qcg_rhf_gradient(qcg_context* ctx, const double* D, const double* C, const double* eps, int n_occ, double* grad);
/* XC part of the nuclear gradient, B3LYP on the context's grid, no grid-weight response: replaces
 * xc_gradient_analytical (scf/gradient.rs:1687-1910, with evaluate_basis_hessian_on_grid :1938).
 * grad (natom x 3) is overwritten. */
int qcg_xc_gradient(qcg_context* ctx, const double* D, double* grad);
/* B3LYP (closed-shell KS) nuclear gradient at a converged KS state: nuclear repulsion + one-electron
 * (D, W) + two-electron with exact-exchange fraction 0.20 + XC, i.e. the gradient assembly of
 * ks_dft_gradient_with_guess (scf/gradient.rs:1214-1341) without D3(BJ).  C: MOs in columns,
 * row-major; eps ascending; grad (natom x 3) overwritten. */
int qcg_ks_gradient(qcg_context* ctx, const double* D, const double* C, const double* eps, int n_occ, double* grad);

#ifdef __cplusplus
}
#endif
#endif /* QCG_H */
