/*
 * lsmsx.h -- C ABI of the MI355X tau00 path: KKR matrix assembly
 * M = 1 - t G(E) over a local interaction zone and the central-block
 * inversion tau00 = [M^-1 [t0; 0]]_00, on one GPU.
 *
 * Plain pointers and sizes only.  Complex numbers are interleaved
 * (re, im) doubles, the layout of std::complex<double>; matrices are
 * column-major (element (i,j) at a[i + j*ld]).  Host buffers are
 * caller-owned; device memory is owned by the context.
 *
 * Return codes: 0 ok; > 0 a LAPACK-style singular-pivot position; < 0 an
 * argument or HIP error (lsmsx_last_error() describes it).  No entry point
 * exits the process.  One context serves one host thread at a time.
 */
#ifndef LSMSX_H
#define LSMSX_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lsmsx_ctx lsmsx_ctx;

/* precision modes for the solve */
#define LSMSX_PREC_FP64 0       /* all arithmetic complex FP64 */
#define LSMSX_PREC_MIXED_FP32 1 /* complex FP32 factors (FP32 MFMA) + FP64 iterative
                                   refinement against the FP64 matrix; solves whose
                                   refinement stalls are redone in FP64 */
#define LSMSX_PREC_MIXED_F16 2  /* as LSMSX_PREC_MIXED_FP32, with the trailing updates of
                                   the FP32 factorization on the FP16 matrix cores
                                   (hi/lo split panels, FP32 accumulation) */
#define LSMSX_PREC_FP64_REFINED 3 /* FP64 factors of a copy of M + FP64 iterative
                                   refinement against M: full FP64 accuracy when the
                                   tile-local pivots are small.  Used automatically for
                                   the plain path's guarded solves and the mixed path's
                                   unconverged solves */

/* negative per-solve info values of the batched calls (see lsmsx_tau00_batched) */
#define LSMSX_INFO_SWEEP_STALL -2
#define LSMSX_INFO_REFINE_UNCONVERGED -3

/* linear-solver selector ids for the reference's 16-bit
 * lsms.global.linearSolver field (linearSolvers.hpp:18-27,196-199;
 * buildKKRMatrix.hpp:22-30): solver bits 0x0fff, build bits 0xf000. */
#define LSMSX_LINEAR_SOLVER_TAU00_MI355X 0x040
#define LSMSX_BUILD_KKR_MATRIX_MI355X 0x4000

/* Last error message of the calling thread ("" if none). */
const char *lsmsx_last_error(void);

/* Create / destroy a context on HIP device `device`.  Owns all device
 * buffers, like DeviceStorage::allocate (Accelerator/DeviceStorage_HIP.cpp:123-233). */
int lsmsx_ctx_create(int device, lsmsx_ctx **out);
int lsmsx_ctx_destroy(lsmsx_ctx *ctx);

/* Angular-momentum tables for maxlmax <= 3: Gaunt numbers cgnt with dims
 * (lmax+1, kkr, kkr), illp(kkr, kkr) = i^(l-l'), ilp1(0..2lmax) = i^(l+1).
 * Replaces DeviceConstants::allocate (Accelerator/DeviceStorage_HIP.cpp:402-)
 * fed from GauntCoeficients::cgnt / IFactors (Misc/Coeficients.hpp:42-106). */
int lsmsx_set_tables(lsmsx_ctx *ctx, int maxlmax, const double *cgnt,
                     const double *illp, const double *ilp1);

/* Compute the same tables on the host (independent of any caller). */
int lsmsx_compute_tables(int maxlmax, double *cgnt, double *illp, double *ilp1);

/* Upload the t-matrix store: tmatStore(r, s) at tmat[2*(r + s*ldTmatStore)],
 * blocks of blkSizeTmatStore complex per energy slot, groupSize slots.
 * Replaces DeviceStorage::copyTmatStoreToDevice (DeviceStorage_HIP.cpp:282-309). */
int lsmsx_upload_tmats(lsmsx_ctx *ctx, const double *tmat, long ldTmatStore,
                       int numStore, long blkSizeTmatStore, int groupSize);

/* Relativistic store (relativity = full): tmat holds the reference's
 * kappa-mu t-matrices, (2 kkrsz)^2 per energy slot, and energies[2 e],
 * energies[2 e + 1] the complex energy of slot e.  The KKR blocks are then
 * M_ij = -t_i (prel^2/E) relmtrx(G_ij) (setgij.f:39-55, relmtrx.f): the
 * batched and single tau00 calls (n_spin_cant = n_spin_pola = 2, every LIZ
 * site at the central lmax) return the kappa-mu tau00 = [M^-1 [t0;0]]_00,
 * which the caller rotates with dmatp tau00 dmat (tau_inv_postproc.f:79-130;
 * no tau00 - t), and lsmsx_build_kkr dumps the kappa-mu M.  Computed as the
 * orthogonal similarity M = U~ M' U~^T of the spin-canted matrix of
 * t' = (prel^2/E) U^T t U (U: the Clebsch-Gordan matrix of clebsch.f), so
 * every precision mode applies.  Full tau and the adjoint route reject it.
 * A later lsmsx_upload_tmats returns to the non-relativistic blocks.
 * Replaces the relativistic branch of the F77 build (calculateTauMatrix.cpp:
 * 216-242) and of the old solvers' post-processing (:818-821). */
int lsmsx_upload_tmats_rel(lsmsx_ctx *ctx, const double *tmat, long ldTmatStore,
                           int numStore, long blkSizeTmatStore, int groupSize,
                           const double *energies);

/* Register the LIZ of local atom `atom_id`: LIZPos (3 x numLIZ, relative to
 * the central site 0), LIZlmax, LIZStoreIdx (AtomData fields,
 * SingleSite/AtomData.hpp:441-467).  Replaces DeviceAtom::copyFromAtom
 * (DeviceStorage_HIP.cpp:380-390). */
int lsmsx_set_liz(lsmsx_ctx *ctx, int atom_id, int numLIZ, const double *LIZPos,
                  const int *LIZlmax, const int *LIZStoreIdx);

/* Assemble the KKR matrix of one atom/energy in the reference's site order
 * into host memory m (nrmat_ns x nrmat_ns, leading dimension ldm).
 * Replaces buildKKRMatrix / buildKKRMatrixCPU / buildKKRMatrixHip
 * (MultipleScattering/buildKKRMatrix.hpp:27-42; calculateTauMatrix.cpp:57). */
int lsmsx_build_kkr(lsmsx_ctx *ctx, int atom_id, int ispin, int iie,
                    const double *prel, int n_spin_cant, int n_spin_pola,
                    double *m, int ldm);

/* Solver-level tau00 = [M^-1 [t0;0]]_00 (kkrsz_ns x kkrsz_ns) for one
 * atom/energy, assembling M on the device.  Replaces the
 * calculateTauMatrix build+solve pair for solver ids < 0xf00
 * (calculateTauMatrix.cpp:359-427,619-686; solveTau00zgetrf,
 * linearSolvers_CPU.cpp:98-123).  The caller then applies
 * calculateTau00MinusT and the local-frame rotation as today. */
int lsmsx_tau00(lsmsx_ctx *ctx, int atom_id, int ispin, int iie,
                const double *prel, int n_spin_cant, int n_spin_pola,
                int precision, double *tau00, int *info);

/* Batched form: nsolve independent (atom, spin, energy slot) solves in one
 * grouped launch sequence.  prel holds 2*nsolve doubles; tau00 receives
 * nsolve blocks of kkrsz_ns^2 complex; info nsolve ints (0 ok, > 0 M is
 * singular: the first exactly singular U(i,i) of an LU with partial
 * pivoting over whole columns, 1-based in the reference order; -2 an
 * internal wait bound of the mixed path's refinement sweep was hit: never
 * expected, the solve's tau00 is then not valid; -3 the FP64 refinement of
 * the solve did not converge within the sweep limit and tau00 is the
 * unrefined plain FP64 result -- only for relativistic stores, see below).
 * Solves that pivoting inside the 64 x 64 diagonal tiles cannot factor (a
 * singular diagonal tile of a non-singular M), and solves whose FP64
 * refinement (LSMSX_PREC_FP64_REFINED, or the redo of a guarded plain solve
 * or of an unconverged mixed solve) does not converge, are solved again by
 * the reference's own route, LU with partial pivoting over whole columns of
 * the FP64 M and M X = [t0; 0] (rocSOLVER zgetrf/zgetrs on the device);
 * lsmsx_last_batch_stats marks them with iters = -2.  Relativistic stores
 * keep the -3 result.  Replaces the OpenMP loop of calculateAllTauMatrices
 * (calculateTauMatrix.cpp:841-945). */
int lsmsx_tau00_batched(lsmsx_ctx *ctx, int nsolve, const int *atom_ids,
                        const int *ispins, const int *iies, const double *prel,
                        int n_spin_cant, int n_spin_pola, int precision,
                        double *tau00, int *info);

/* tau00 from a caller-built matrix m (n x n, central kkrsz_ns block first)
 * and t0 (k x k).  m is not modified.  Same contract as solveTau00zgetrf
 * (linearSolvers_CPU.cpp:98-123) with the matrix on the host.  Pivoting is
 * inside 64-row diagonal tiles; when a tile meets a zero pivot the solve is
 * repeated once with the non-central rows in a fixed pseudo-random order
 * (tau00 does not change under it), and info reports the first attempt's
 * pivot only if that also fails. */
int lsmsx_solve_tau00_matrix(lsmsx_ctx *ctx, int n, int k, const double *m,
                             int ldm, const double *t0, double *tau00, int *info);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): lsmsx_solve_tau00_matrix with every buffer on the
 * device -- d_m (n x n, leading dimension ldm, central block first, not
 * modified), d_t0 (k x k), d_tau00 (k x k, written) and d_info (one int,
 * written: 0 or the first singular pivot) -- and all scratch in a fixed
 * per-GPU static device array: no hipMalloc, hipFree or allocating library
 * call.  Limit: roundup(n, 64) <= 1024 (returns -9 beyond it).  The array is
 * shared by every call on one GPU and keeps the last call's contents: calls
 * that overlap on one GPU (several host threads, or several contexts of one
 * device) race and must be serialized by the caller.  Runs on the context's
 * stream and returns once the solve is done. */
int lsmsx_solve_tau00_matrix_noalloc(lsmsx_ctx *ctx, int n, int k, const double *d_m,
                                     int ldm, const double *d_t0, double *d_tau00,
                                     int *d_info);

/* Refinement settings of LSMSX_PREC_MIXED_FP32: at most max_sweeps
 * residual/correction sweeps (default 12), converged when the central rows
 * satisfy ||dX||_F <= tol ||X||_F (default 1e-12).  Solves that do not
 * converge are redone in FP64. */
int lsmsx_set_refinement(lsmsx_ctx *ctx, int max_sweeps, double tol);

/* Residual form of the mixed modes (no reference counterpart):
 *   LSMSX_RESIDUAL_AUTO    (default) when every solve of a batch is collinear
 *                          with 16 right-hand sides and every LIZ site has
 *                          lmax 3, the assembly stores the structure
 *                          constants G_ij once per site pair (G_ji =
 *                          (-1)^(l1+l2) G_ij entry by entry) instead of the
 *                          FP64 M, and each residual is formed as
 *                          B - X + t (G X) from it: half the bytes of M;
 *                          other batches use M;
 *   LSMSX_RESIDUAL_MATRIX  always the FP64 M.
 * The two forms round differently; both refine to the same tolerance. */
#define LSMSX_RESIDUAL_AUTO 0
#define LSMSX_RESIDUAL_MATRIX 1
int lsmsx_set_residual_form(lsmsx_ctx *ctx, int form);

/* Pivot guard of the plain FP64 path: a solve whose smallest tile-local pivot
 * (|u_kk|, pivoting inside 64x64 diagonal tiles) is below threshold is redone
 * with LSMSX_PREC_FP64_REFINED (default 1e-2; 0 disables).  Small pivots mean
 * element growth, which can cost tau00 digits on ill-conditioned matrices;
 * the refinement against the untouched M recovers them.  No counterpart in the
 * reference (LAPACK zgetrf pivots over whole columns). */
int lsmsx_set_pivot_guard(lsmsx_ctx *ctx, double threshold);

/* Per-solve statistics of the last batched / single call on this context:
 * iters[i] = refinement sweeps (0 on the plain FP64 path, > 0 for a refined
 * solve including the plain path's guarded ones, -1 when a mixed solve was
 * redone in FP64, -2 when the column-pivoted fallback produced the result),
 * minpiv[i] = smallest |pivot| of the tile-pivoted factorization of the
 * solve.  Either pointer may be NULL. */
int lsmsx_last_batch_stats(lsmsx_ctx *ctx, int nsolve, int *iters, double *minpiv);

/* Workspace cap in bytes for the batched matrix pool (0 = automatic:
 * a fraction of the free device memory).  Larger batches run in chunks. */
int lsmsx_set_workspace_limit(lsmsx_ctx *ctx, long bytes);

/* Kernel timing: when enabled, every launch is bracketed by HIP events on
 * the context's stream; lsmsx_get_profile returns the accumulated time (ms)
 * and launch count of one kernel kind (0 assemble, 1 diagonal-tile LU,
 * 2 panel, 3 trailing update, 4 last tile, 5 refinement residual,
 * 6 refinement solve, 7 explicit inverse, 8 Lloyd/adjoint contraction,
 * 9 cross update of the paired-step elimination).
 * lsmsx_reset_profile clears them. */
int lsmsx_set_profiling(lsmsx_ctx *ctx, int on);
int lsmsx_reset_profile(lsmsx_ctx *ctx);
int lsmsx_get_profile(lsmsx_ctx *ctx, int kind, double *ms, long *launches);

/* Upload dt/dE for every entry of the t-matrix store, in the store's layout
 * and dimensions (needed for the Lloyd term).  The reference has no such
 * store: its t-matrices come from the single-site solver at each energy
 * (SingleSite/SingleSiteScattering.cpp:6-54); the caller supplies their
 * energy derivative. */
int lsmsx_upload_dtmats(lsmsx_ctx *ctx, const double *dtmat, long ldTmatStore,
                        int numStore, long blkSizeTmatStore, int groupSize);

/* Full tau = M^-1 T (nrmat_ns x nrmat_ns, reference site order, leading
 * dimension ldtau) with T the block-diagonal of the LIZ sites' t-matrices
 * of energy slot iie, from an explicit FP64 inverse of the device-built M.
 * Replaces solveTauFullzgetrf (linearSolvers_CPU.cpp:125-144, zgetrf +
 * zgetrs with N right-hand sides; the reference always reads slot 0 and
 * assumes identical lmax: pass iie = 0 to match it, reduced-lmax sites use
 * the packing of the assembly) as called by Kubo/CurrentMatrix.cpp:184-190. */
int lsmsx_tau_full(lsmsx_ctx *ctx, int atom_id, int ispin, int iie,
                   const double *prel, int n_spin_cant, int n_spin_pola,
                   double *tau, int ldtau, int *info);

/* Adjoint route (FP64, explicit inverse of each M from its LU).  For nsolve
 * (atom, spin, slot) solves with energies E (2*nsolve doubles, needed when
 * lloyd != NULL) and prel = sqrt(E (1 + E/c^2)) (2*nsolve doubles, != 0):
 *   tau00[b]  (kkrsz_ns^2 complex, may be NULL) as lsmsx_tau00_batched;
 *   lloyd[b]  (1 complex, may be NULL) = Tr(M^-1 dM/dE) = d ln det M / dE,
 *             dM/dE = -(dt/dE G + t dG/dE), dt/dE from lsmsx_upload_dtmats;
 *   with tau00_bar (nsolve kkrsz_ns^2 complex cotangents, may be NULL):
 *   tmat_grad (ldTmatStore x numStore complex, overwritten) = the vector-
 *             Jacobian product sum_b J_b^H tau00_bar[b] of tau00 with respect
 *             to every t-matrix store entry (conjugate convention of complex
 *             reverse-mode differentiation: for tau00 = R t C, grad_t =
 *             R^H bar C^H).
 * kkrsz_ns must be 16 or 32.  No counterpart in the reference (it has no
 * Lloyd term or derivative of tau); the tau00 it returns agrees with
 * lsmsx_tau00_batched (solves under the pivot guard take their tau00 from
 * LSMSX_PREC_FP64_REFINED). */
int lsmsx_tau00_adjoint(lsmsx_ctx *ctx, int nsolve, const int *atom_ids,
                        const int *ispins, const int *iies, const double *energies,
                        const double *prel, int n_spin_cant, int n_spin_pola,
                        double *tau00, double *lloyd, const double *tau00_bar,
                        double *tmat_grad, int *info);

/* Schedule of the batched factorization (no effect on results, which are
 * bitwise identical under every schedule):
 *   LSMSX_SCHED_RIGHT_ALIGN  solves of different size start their elimination
 *                            late enough that all finish on the same step, so
 *                            the small final steps of every solve run together;
 *   LSMSX_SCHED_XCD_SWIZZLE  trailing-update workgroups (and, with split-FP16
 *                            updates, the panels with their fused cross
 *                            updates) are remapped so that each of the 8 XCDs
 *                            works on whole solves and reads their panels
 *                            into its own L2;
 *   LSMSX_SCHED_TWO_STREAMS  the two halves of a chunk run on two streams (of
 *                            default priority), so one half's diagonal-tile
 *                            and panel steps overlap the other's updates.
 * No reference counterpart (the reference solves one matrix per OpenMP
 * thread, calculateTauMatrix.cpp:891-900). */
#define LSMSX_SCHED_RIGHT_ALIGN 1
#define LSMSX_SCHED_XCD_SWIZZLE 2
#define LSMSX_SCHED_TWO_STREAMS 4
int lsmsx_set_schedule(lsmsx_ctx *ctx, int flags);
int lsmsx_get_schedule(lsmsx_ctx *ctx, int *flags);

/* hipGraph replay of the batched path: when on, the assembly + factorization
 * launch sequence of each chunk shape is captured once and relaunched as one
 * graph on later calls with the same shape (the refinement sweeps of the
 * mixed path, which wait on convergence counts, stay outside the graph). */
int lsmsx_set_graph(lsmsx_ctx *ctx, int on);

/* Structure constants shared within a batch (on by default): solves of one
 * chunk whose atoms have bitwise equal LIZ positions and LIZlmax (atoms of a
 * lattice whose LIZs are built alike) and the same energy slot and prel
 * share G_ij(E); the first of them evaluates it once per site pair and every
 * solve builds its blocks M_ij = -t_i G_ij from the stored values, so the
 * matrices (and results) are bitwise those of per-solve evaluation.
 * Collinear, non-relativistic batches; off: every solve evaluates its own.
 * No reference counterpart (the reference builds G per atom and energy,
 * calculateTauMatrix.cpp:57-315). */
int lsmsx_set_shared_structure_constants(lsmsx_ctx *ctx, int on);

/* Steps per trailing update of the split-FP16 path (LSMSX_PREC_MIXED_F16):
 * 4, 8 (the default) or 16.  Larger groups read and write the trailing
 * matrix less often and spend more in the cross updates inside the panels,
 * which apply up to group - 1 earlier steps of the group to each panel
 * tile; results of different group sizes differ by rounding only
 * (each meets the refinement tolerance).  lsmsx_get_split_group returns the
 * current size and, in ms[3] (may be NULL), the wall times of sizes 4, 8 and
 * 16 measured by the last lsmsx_autotune at LSMSX_PREC_MIXED_F16 (-1 when
 * not measured).  No reference counterpart (tuning of this path's
 * factorization). */
int lsmsx_set_split_group(lsmsx_ctx *ctx, int group);
int lsmsx_get_split_group(lsmsx_ctx *ctx, int *group, double *ms);

/* Autotuner: times the given batch (arguments as lsmsx_tau00_batched) under
 * each of the eight schedules (flag combinations), keeps the fastest on the
 * context and returns it in *best; ms[8] (may be NULL) receives the wall
 * times.  At LSMSX_PREC_MIXED_F16 it then times the split group sizes 4, 8
 * and 16 under that schedule and keeps the fastest (lsmsx_get_split_group),
 * or 8 when that is within 2% of the fastest (one timed run per size).
 * The tile width (64) and the update kernels' region shapes are fixed: the
 * MFMA wave tiles, the LDS-resident diagonal LU and the tile-local pivoting
 * are built on them. */
int lsmsx_autotune(lsmsx_ctx *ctx, int nsolve, const int *atom_ids, const int *ispins,
                   const int *iies, const double *prel, int n_spin_cant, int n_spin_pola,
                   int precision, int *best, double *ms);

/* Single-site scattering of npot spherical potentials at ne energies:
 * calculateSingleScattererSolution (SingleSite/SingleSiteScattering.cpp:6-54)
 * calls, per spin, single_scatterer_nonrel_ (SingleSite/single_scatterer_nonrel.f:4-122),
 * which solves the radial equation through semrel (SingleSite/semrel_c.f:3-228,
 * solver 2: Misc/fitpot.f, Misc/rwave.f).  One entry for the whole batch; a
 * spin-polarized atom passes each spin's potential as its own p.
 *   nrelv           0: scalar-relativistic, else non-relativistic (lsms.nrelv)
 *   lmax            1..3 (atom.lmax)
 *   vr, r_mesh      npot x iprpts: r v(r) and the radial mesh of potential p at
 *                   [p*iprpts + j] (vr(0,is), atom.r_mesh); r_mesh[0] > 0, increasing
 *   jmt, jws        1-based mesh indices of the muffin-tin and Wigner-Seitz radii;
 *                   max(jmt+11, jws) <= iprpts (semrel_c.f:131-143)
 *   r_sph           matching radius (atom.rInscribed, or atom.rws when mtasa > 0)
 *   energy, prel    ne complex: E and sqrt(E (1 + E/c^2))
 * Outputs (host; q = e*npot + p):
 *   matom           [(q*(lmax+1) + l)] complex, 1/t_l (solution.matom(l, is))
 *   tmat            may be NULL; [q*kkrsz*kkrsz + ...] column-major kkrsz x kkrsz,
 *                   diagonal 1/matom_l with kkrsz = (lmax+1)^2 (solution.tmat_l)
 *   zlr, jlr        may be NULL; [(q*(lmax+1) + l)*iprpts + j] complex for j < jws
 *                   (solution.zlr, solution.jlr; other j are zero)
 * Returns 0; 1 if an adaptive integration step underflowed somewhere (the
 * reference prints a warning and goes on; the results are still written).
 * The inward (irregular) integration feeds only jlr and runs only when jlr
 * is requested;
 * < 0 on bad arguments.  The spin-canted frame rotation (trltog) is done
 * by lsmsx_tmats_from_single_site_canted when the store is filled on the
 * device. */
int lsmsx_single_site_nonrel(lsmsx_ctx *ctx, int nrelv, int lmax, int npot, int iprpts,
                             const double *vr, const double *r_mesh, const int *jmt,
                             const int *jws, const double *r_sph, int ne, const double *energy,
                             const double *prel, double *matom, double *tmat, double *zlr,
                             double *jlr);

/* Muffin-tin density of states and spherical Green function of nsolve
 * (atom, energy) pairs of one collinear spin channel, from tau00 and the
 * single-site radial functions: the sphere part of green_function_
 * (MultipleScattering/green_function.f:107-124 for isp = 1, gf_local.f:109-189,
 * int_zz_zj.f:116-165, mdosms_c.f, mgreen_c.f), which is the whole-cell
 * result for ASA potentials (mtasa = 1, int_zz_zj.f:226-237).  For
 * muffin-tin potentials lsmsx_green_interstitial adds the interstitial part
 * (inter_m) of the cell.
 *   pot_of[b]       the potential (0..npot-1) whose mesh solve b uses
 *   r_mesh, jmt, jws, r_sph   per potential, as for lsmsx_single_site_nonrel
 *   tau00           nsolve x kkrsz x kkrsz complex, column-major, local frame
 *                   (tau00_l), kkrsz = (lmax+1)^2
 *   zlr, jlr        nsolve x (lmax+1) x iprpts complex (the single-site layout)
 * Outputs: dos[b] (complex; -Im/pi convention of mdosms, times 2 when
 * n_spin_pola = 1, as green_function.f:184-191 does), green (may be NULL):
 * [b*iprpts + j] for j < jws, zero beyond. */
int lsmsx_green_mt(lsmsx_ctx *ctx, int nsolve, int lmax, int iprpts, int n_spin_pola, int npot,
                   const int *pot_of, const double *r_mesh, const int *jmt, const int *jws,
                   const double *r_sph, const double *tau00, const double *zlr,
                   const double *jlr, double *dos, double *green);

/* Interstitial part of the cell integrals of muffin-tin potentials and its
 * density of states, for nsolve (atom, energy) pairs: inter_m /
 * interstitial (VORPOL/inter.f, VORPOL/interstitial.f), which int_zz_zj adds
 * to the sphere part when mtasa != 1 (int_zz_zj.f:230-252).  The whole-cell
 * DOS is the sum of lsmsx_green_mt's dos and this one (mdosms is linear in
 * pzz and pzj).  Per solve b: prel[b] (complex), matom_left/right[b][l]
 * (complex, l = 0..lmax; equal for collinear spins), tau00[b] (kkrsz x
 * kkrsz complex, column-major, local frame), cell_of[b].  Per cell k (as
 * VORPOL passes them): r_sph[k], ncrit[k] (<= maxreg + 1), grwylm and
 * gwwylm [k][ir][ng] (radial Gauss points and weights, ngaussr per region,
 * maxreg regions per cell, points with r <= 0 skipped), wylm [k][ir][ng][i]
 * (complex step-function coefficients, i = l(l+1)/2 + m, m >= 0, l <=
 * 2 lmax).  The Gaunt numbers are the uploaded tables (lmax <= their lmax).
 * Outputs: dos[b] (complex, -Im/pi convention of mdosms, times 2 when
 * n_spin_pola = 1 as for lsmsx_green_mt), and if not NULL pzz[b] (kkrsz x
 * kkrsz, pzz(L2, L1) at [L2 + kkrsz L1]) and pzj[b] (complex). */
int lsmsx_green_interstitial(lsmsx_ctx *ctx, int nsolve, int lmax, int zj_flag,
                             int n_spin_pola, const double *prel, const double *matom_left,
                             const double *matom_right, int ncell, const int *cell_of,
                             const double *r_sph, const int *ncrit, int ngaussr, int maxreg,
                             const double *grwylm, const double *gwwylm, const double *wylm,
                             const double *tau00, double *pzz, double *pzj, double *dos);

/* Keep the radial functions of the next lsmsx_single_site_nonrel calls on
 * the device (computed even when zlr/jlr are NULL) for
 * lsmsx_green_mt_resident.  No reference counterpart: the reference keeps
 * zlr/jlr in host memory between the single-site solve and green_function_
 * (energyContourIntegration.cpp:426-436 and :527-546); this pair keeps the
 * chain in HBM. */
int lsmsx_set_single_site_resident(lsmsx_ctx *ctx, int on);

/* As lsmsx_green_mt, with the radial functions, meshes and lmax of the last
 * lsmsx_single_site_nonrel call made with residency on: solve b uses the
 * (energy e, potential p) pair pair_of[b] = e*npot + p of that call.
 * jmt_of_pot (may be NULL: the single-site jmt) sets the end of the sphere
 * integrals per potential; the reference passes jws there for ASA
 * (energyContourIntegration.cpp:519).
 * Returns -3 when no such call precedes it (an lsmsx_green_mt call in between
 * reuses the buffers and also counts as none).  lmax and iprpts size the
 * caller's tau00 (nsolve x kkrsz^2) and green (nsolve x iprpts) buffers and
 * must equal those of that call (-2 otherwise).  No reference counterpart
 * (see lsmsx_set_single_site_resident). */
int lsmsx_green_mt_resident(lsmsx_ctx *ctx, int nsolve, int lmax, int iprpts, int n_spin_pola,
                            const int *pair_of, const int *jmt_of_pot, const double *tau00,
                            double *dos, double *green);

/* Fill the device t-matrix store from the last lsmsx_single_site_nonrel
 * batch without a host round trip: the store (same dimensions as
 * lsmsx_upload_tmats) is zeroed, as local.tmatStore = 0 before the
 * single-site solve (energyContourIntegration.cpp:420), then energy e of
 * potential p goes to slot e of column store_of_pot[p], spin block
 * spin_of_pot[p] (offset spin*kkrsz^2, the collinear layout), as
 * solveSingleScatterers fills &local.tmatStore(iie*blkSizeTmatStore, i)
 * (Main/solveSingleScatterers.cpp:26-28, SingleSiteScattering.cpp:28-49).
 * The spin-canted global-frame t (trltog) is not formed here. */
int lsmsx_tmats_from_single_site(lsmsx_ctx *ctx, long ldTmatStore, int numStore,
                                 long blkSizeTmatStore, int groupSize, const int *store_of_pot,
                                 const int *spin_of_pot);

/* Spin-canted muffin-tin DOS and Green function: green_function_ with
 * n_spin_cant = 2 (MultipleScattering/green_function.f:84-180), sphere part
 * as lsmsx_green_mt: the four spin channels (zsl, zsr) = (1,1), (2,1),
 * (1,2), (2,2) of gf_local, the irregular term on the diagonal ones only,
 * then Tr and the x, y, z parts with wx, wy, wz (Misc/u_sigma_u.f).
 *   tau00_l   nsolve x 4 x kkrsz x kkrsz complex: the four column-major
 *             blocks trgtol.f writes (ml(nr,nc,4))
 *   zlr, jlr  nsolve x 2 spins x (lmax+1) x iprpts complex
 *   wxyz      nsolve x 12 complex: wx[4], wy[4], wz[4] of the solve's atom
 * Outputs: dos[b*4 + a] (a = Tr, x, y, z), green (may be NULL)
 * [(b*4 + a)*iprpts + j] for j < jws. */
int lsmsx_green_mt_canted(lsmsx_ctx *ctx, int nsolve, int lmax, int iprpts, int npot,
                          const int *pot_of, const double *r_mesh, const int *jmt,
                          const int *jws, const double *r_sph, const double *tau00_l,
                          const double *zlr, const double *jlr, const double *wxyz,
                          double *dos, double *green);

/* Spin-canted variant of lsmsx_tmats_from_single_site: column
 * store_of_atom[a] gets, at every energy slot e, the global-frame
 * 2kkrsz x 2kkrsz t = trltog(t_up, t_dn) (Misc/trltog.f:1-52, as
 * calculateSingleScattererSolution does, SingleSite/SingleSiteScattering.cpp:44-46)
 * of the atom's spin-up and spin-down potentials pot_up[a], pot_dn[a] of the
 * last single-site batch, with the atom's ubr, ubrd (4 complex each,
 * AtomData::ubr/ubrd) at [8a .. 8a+7] doubles; leading dimension 2kkrsz.
 * The store is zeroed first. */
int lsmsx_tmats_from_single_site_canted(lsmsx_ctx *ctx, long ldTmatStore, int numStore,
                                        long blkSizeTmatStore, int groupSize, int natom,
                                        const int *store_of_atom, const int *pot_up,
                                        const int *pot_dn, const double *ubr, const double *ubrd);

#ifdef __cplusplus
}
#endif
#endif
