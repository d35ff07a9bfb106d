/*
 * ofdft_hip.h -- C ABI of libofdft_hip.so, the MI355X (gfx950) implementation of
 * the orbital-free DFT energy/potential evaluation, density optimizer and
 * Crank-Nicolson propagator of DFTpy (shaoxc/dftpy).
 *
 * Conventions shared by every entry point
 *  - Device buffers are owned by the caller; pointers are plain device
 *    addresses, C-contiguous, real fields (nx,ny,nz) and R2C spectra
 *    (nx,ny,nz/2+1) with the last axis contiguous, full complex fields
 *    (nx,ny,nz).  Real type is double (precision OFDFT_FP64) or float
 *    (OFDFT_FP32); complex is the matching interleaved (re,im) pair.
 *  - Scalar results are FP64 in both precisions and are written to device
 *    memory (double*), so a call never blocks the host.
 *  - `stream` is a hipStream_t (NULL = default stream).  Calls on one plan
 *    must be ordered on one stream; different plans are independent.
 *  - Return value: 0 on success, nonzero on error; ofdft_last_error() gives
 *    the message (thread-local).
 *  - The plan owns its workspace (allocated at create, freed at destroy);
 *    the FFT, evaluation, reduction, optimizer and propagator entry points
 *    allocate no device memory.  Setup calls allocate: plan-owned tables
 *    that grow only (set_lmgp, set_tn_precond, plan_reserve_lbfgs, the PP
 *    staging and PME moduli, the force/stress workspace) and the per-call
 *    upload of ofdft_atomic_density.
 *
 * Each declaration names the reference interface it replaces, as
 * path:line relative to the reference's src/dftpy.
 */
#ifndef OFDFT_HIP_H
#define OFDFT_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ofdft_plan ofdft_plan;

enum { OFDFT_FP64 = 0, OFDFT_FP32 = 1 };

/* calcType flags of AbstractFunctional.__call__ (functional/abstract_functional.py:22-28) */
enum { OFDFT_CALC_E = 1, OFDFT_CALC_V = 2 };

/* functional terms of a TotalFunctional (functional/total_functional.py:49-59) */
enum {
    OFDFT_TERM_TF = 1,      /* functional/kedf/tf.py:65-80 */
    OFDFT_TERM_VW = 2,      /* functional/kedf/vw.py:114-131 */
    OFDFT_TERM_WT = 4,      /* functional/kedf/wt.py:108-132 */
    OFDFT_TERM_XC = 8,      /* functional/xc/semilocal_xc.py:252-290 (PZ LDA) */
    OFDFT_TERM_HARTREE = 16,/* functional/hartree.py:22-45 */
    OFDFT_TERM_PP = 32,     /* functional/pseudo/__init__.py:219-234 */
    OFDFT_TERM_LMGP = 64    /* functional/kedf/lwt.py:501-575 (LMGP-NL), needs ofdft_set_lmgp */
};

/* energies[] layout written by ofdft_eval (Ha) */
enum { OFDFT_E_TF = 0, OFDFT_E_VW, OFDFT_E_NL, OFDFT_E_XC, OFDFT_E_H, OFDFT_E_PP, OFDFT_E_SUM, OFDFT_E_NELEC, OFDFT_E_COUNT };

/* Library version and last error message (thread-local). */
int ofdft_version(void);
const char *ofdft_last_error(void);

/* Plan = DirectGrid + ReciprocalGrid + FFT plans + workspace.
 * Replaces DirectGrid(lattice, nr, full) (grid.py:302-318) and
 * get_reciprocal() (grid.py:384-415).  lattice: rows are the cell vectors
 * in bohr.  full != 0 reproduces the reference's full-complex-grid results
 * (fftfreq on all axes) exactly on the R2C half grid.  max_batch bounds
 * the batch argument of the FFT calls (workspace sizing). */
int ofdft_plan_create(const int nr[3], const double lattice[9], int precision, int full,
                      int max_batch, ofdft_plan **plan);
int ofdft_plan_destroy(ofdft_plan *plan);
/* Replace the cell of a plan (same grid shape): a strained copy of the cell
 * reuses the FFT plans and workspace (Ions/DirectGrid rebuilt on a new cell
 * in the reference, grid.py:302-318). */
int ofdft_plan_set_lattice(ofdft_plan *plan, const double lattice[9]);
/* FFT kernel selection (no counterpart in the reference): fixed != 0 (the
 * default) uses the size-specialised kernels for the lengths that have them,
 * 0 forces the generic mixed-radix kernels everywhere. */
int ofdft_plan_set_fft_kernels(ofdft_plan *plan, int fixed);
/* FFT autotuner (no counterpart in the reference): times the variants of each
 * size-specialised pass on the plan's shapes and keeps the fastest.  The z
 * R2C and C2R row passes have 3 (0..2: default tile, halved, quartered); the
 * x and y column passes 6: tile v % 3 with the radix order v / 3 (0: radix
 * 8s first, as the length factorises; 1: reversed, odd radices first).
 * variants (optional, 4 ints): chosen variant of the z R2C, z C2R, x and y
 * passes; times (optional, 24 floats): mean ms of each variant, 6 per pass
 * in that order (-1 where a pass has fewer).  The choice for a grid shape and
 * precision is kept for the process: later plans of that shape take it
 * without timing (times all -1), so every plan of a shape gives identical
 * transforms.  Synchronises stream. */
int ofdft_plan_autotune(ofdft_plan *plan, int *variants, float *times, void *stream);
/* Set the tile variants explicitly (same order as above). */
int ofdft_plan_set_variants(ofdft_plan *plan, const int *variants);
/* Bytes of device workspace the plan holds. */
size_t ofdft_plan_workspace_bytes(const ofdft_plan *plan);

/* FFT with DFTpy's normalisation.
 * r2c: DirectField.fft = rfftn * dV (field.py:373-390), batch fields.
 * c2r: ReciprocalField.ifft = irfftn(s=nr) / dV (field.py:739-764).
 * c2c: full complex grid, forward fftn * dV / inverse ifftn / dV
 *      (field.py:159-172 cplx branch).  In-place allowed for c2c only. */
int ofdft_fft_r2c(ofdft_plan *plan, const void *in, void *out, int batch, void *stream);
int ofdft_fft_c2r(ofdft_plan *plan, const void *in, void *out, int batch, void *stream);
int ofdft_fft_c2c(ofdft_plan *plan, const void *in, void *out, int inverse, int batch, void *stream);

/* Functional configuration.
 * KEDF options (functional/kedf/__init__.py:90-184): x scales TF, y scales
 * vW, alpha/beta the WT exponents, rho0 <= 0 means rho0 = mean(rho)
 * (wt.py:112-113), sigma > 0 applies vW's Gaussian filter (vw.py:45-49).
 * The WT kernel (kernel.py:325-334) is evaluated in-kernel from |G|. */
int ofdft_set_kedf(ofdft_plan *plan, double x, double y, double alpha, double beta,
                   double rho0, double sigma);
/* Active terms (OFDFT_TERM_* mask) of the TotalFunctional. */
int ofdft_set_terms(ofdft_plan *plan, unsigned terms);
/* Local pseudopotential in real space (LocalPseudo._vreal,
 * pseudo/__init__.py:236-246): device array of N reals in the plan
 * precision, kept by pointer (caller keeps it alive) or NULL. */
int ofdft_set_vloc(ofdft_plan *plan, const void *vreal);
/* Core density of the non-linear core correction (XC(core_density=...),
 * xc/__init__.py:49-74): the XC term is evaluated at rho + core (energy and
 * potential); device array of N reals in the plan precision, kept by pointer,
 * or NULL.  Not covered by the analytic Hessian product (the forward-difference
 * one applies) nor by the XC stress. */
int ofdft_set_core(ofdft_plan *plan, const void *core);
/* LMGP nonlocal kinetic functional (LWT with the MGP kernel, lwt.py:501-578;
 * OFDFT_TERM_LMGP).  ofdft_mgp_kernel_table computes the MGP kernel table
 * K(eta_k), eta_k = k etamax / (neta - 1) (kernel.py:266-322, t-integral with
 * maxpoints terms) on the GPU into out (host); the inverse Lindhard function
 * inside it is the cubic spline (4 (ng-1) coefficients gcoef) on the
 * log-spaced grid gp[ng], clamped at its ends, as the reference evaluates it.  ofdft_set_lmgp takes the
 * piecewise-cubic tables of K and of -eta dK/deta on that grid (4 (ntab-1)
 * coefficients each, host) and the k_F ladder ratio (1.15 in the reference),
 * and allocates the plan's LMGP workspace (setup call); ntab <= 0 disables.
 * The saved k_F bounds (lwt.py:36-82) persist across evaluations until
 * ofdft_lmgp_reset or the next ofdft_set_lmgp. */
int ofdft_mgp_kernel_table(int neta, double etamax, int maxpoints, int ng, const double *gp, const double *gcoef,
                           double *out);
int ofdft_set_lmgp(ofdft_plan *plan, int ntab, double etamax, const double *kcoef, const double *dcoef,
                   double ratio);
int ofdft_lmgp_reset(ofdft_plan *plan);
/* Ion-ion (Ewald) energy added to OFDFT_E_SUM (total_functional.py:134-137). */
int ofdft_set_eion(ofdft_plan *plan, double eion);

/* One TotalFunctional evaluation (total_functional.py:115-146):
 * energies (device, OFDFT_E_COUNT doubles) and/or the total potential pot
 * (device, N reals).  rho is not modified (unlike wt.py:15-26). */
int ofdft_eval(ofdft_plan *plan, const void *rho, unsigned calc, void *pot, double *energies,
               void *stream);
/* The same evaluation with the calcType 'D' and 'V2' fields of the terms
 * that have them in the reference (device, N reals each, either may be
 * NULL): energy density edens = TF C rho^(5/3) (kedf/tf.py:68-73, without the
 * x factor, as there) + vW y v_vW rho (kedf/vw.py:125-126) + WT rho^alpha
 * (K * rho^beta) (kedf/wt.py:44-56,123-129); v2rho2 = TF x 10/9 C
 * rho^(-1/3) (kedf/tf.py:47-50,77-79) + PZ-LDA d v_xc / d rho
 * (xc/semilocal_xc.py:291-303).  Hartree and the local PP contribute to
 * neither (the reference's return no such fields); LMGP 'D' is refused.
 * Runs the full E+V pipeline (pot may be NULL). */
int ofdft_eval_fields(ofdft_plan *plan, const void *rho, unsigned calc, void *pot, double *energies, void *edens,
                      void *v2rho2, void *stream);

/* Sum_r a(r) b(r) in FP64 (MP.einsum, mpi/mpi.py:199-206); b == NULL gives
 * Sum_r a(r).  Not multiplied by dV. */
int ofdft_dot(ofdft_plan *plan, const void *a, const void *b, double *out, void *stream);

/* Measurement entry point (no counterpart in the reference): launches one
 * pass of the evaluation pipeline `iters` times on `stream` between two HIP
 * events and returns the mean duration in ms.  which: 0 z-pass R2C with the
 * vW/WT prologue (reads rho), 1 y-pass, 2 x-pass (C2C on the half grid),
 * 3 z-pass C2R, 4 full-grid C2C 3D transform, 5 one complete E+V evaluation
 * of rho with the plan's functional settings, 6 the evaluation's x pass
 * (forward C2C, G-space multiply of every spectrum, inverse C2C, in place),
 * 7 the evaluation's point-wise post pass (reads rho); batch 1..3 spectra
 * (passes 0-4).  Uses the plan's workspace as data. */
int ofdft_bench_pass(ofdft_plan *plan, int which, int batch, int iters, const void *rho, double *avg_ms,
                     void *stream);

/* Local PP setup on the grid (LocalPseudo.local_PP, pseudo/__init__.py:236-246,
 * _PP_Reciprocal 377-387, _PP_Reciprocal_PME 389-404).
 * Species s (0..nspecies-1) has a radial table as a piecewise cubic on a
 * uniform grid: q0, dq, ncoef intervals, coef[4*ncoef] (value = c0 + c1 t +
 * c2 t^2 + c3 t^3, t = q - (q0 + i dq)), qmax (zero beyond), and the quartic
 * patch below qsmall: vsmall[3] = (v0, v1, v2), dqs.  Host arrays.
 * Atoms: pos_cart[3*nat] (bohr), pos_frac[3*nat], species[nat].
 * pme != 0 uses order-`order` cardinal B-splines; else the exact
 * structure factor.  Writes vreal (device, N reals, plan precision: per-point
 * math in FP64, the B-spline grid and the transforms in the plan's precision). */
int ofdft_local_pp(ofdft_plan *plan, int nspecies, const double *q0, const double *dq,
                   const int *ncoef, const double *const *coef, const double *qmax,
                   const double *vsmall, const double *dqs, int nat, const double *pos_cart,
                   const double *pos_frac, const int *species, int pme, int order, void *vreal,
                   void *stream);
/* The learned warm start's superposition of screened ions (no counterpart in
 * the reference): as ofdft_local_pp with each species' radial table v_s(q)
 * replaced by rho_s(q) = -v_s(q) / K(q), K the uniform-gas response (resp as
 * in ofdft_set_tn_precond_uniform), evaluated per G point; below the first
 * table step dqs the quartic through (r2, r1, zval, r1, r2) (zval, r1, r2:
 * host arrays of nspecies, the screened values at q = 0, q1, q2).  Writes out
 * (device, N reals, plan precision). */
int ofdft_screened_superposition(ofdft_plan *plan, int nspecies, const double *q0, const double *dq,
                                 const int *ncoef, const double *const *coef, const double *qmax,
                                 const double *vsmall, const double *dqs, const double *zval,
                                 const double *r1, const double *r2, const double *resp, int nat,
                                 const double *pos_cart, const double *pos_frac, const int *species, int pme,
                                 int order, void *out, void *stream);

/* Ewald energy (ewald.py:197-506): real-space erfc sum over images
 * (Energy_real, 280-320), reciprocal sum over the plan's G grid, exact
 * (Energy_rec, 448-465) or PME (Energy_rec_PME, 786-808), and the
 * correction (Energy_corr, 467-482).  eta as in Get_Best_eta (256-279).
 * out (host): [e_total, e_corr, e_real, e_rec, eta].  Synchronises stream. */
int ofdft_ewald(ofdft_plan *plan, int nat, const double *pos_cart, const double *pos_frac,
                const double *charges, int pme, int order, double precision, double *out,
                void *stream);

/* Forces and stress (FP64 plans; host outputs; synchronise stream).  G-space
 * sums run over the half-space mask; these entry points use exact structure
 * factors (the reference's PME=False forms), the *_pme ones below the PME
 * forms.
 *
 * ofdft_local_pp_forces: LocalPseudo.forces (pseudo/__init__.py:456-470),
 *   F_a = (2/V) sum_mask G vloc_s(|G|) Im(rho(G) exp(iG.R_a)); table arguments
 *   as ofdft_local_pp.  forces: nat x 3.
 * ofdft_local_pp_stress: LocalPseudo.stress (pseudo/__init__.py:414-454) with
 *   the spline derivative of vloc; energy = the PP energy at rho.  stress: 3 x 3.
 * ofdft_ewald_forces_stress: ewald.forces / ewald.stress (ewald.py:510-539):
 *   real-space sums over images (Forces_real 541-578, Stress_real 609-662)
 *   plus Forces_rec (580-607) / Stress_rec (719-760) with the G = 0 term;
 *   forces (nat x 3) and stress (3 x 3) each optional.
 * ofdft_functional_stress: out (5 x 9) = stress of TF (kedf/tf.py:52-62), vW
 *   (kedf/vw.py:97-111), WT nonlocal (kedf/wt.py:75-105), PZ LDA
 *   (xc/semilocal_xc.py:308-322) and Hartree (hartree.py:67-91) for the plan's
 *   terms (zero for absent ones); energies (host, OFDFT_E_COUNT) are those of
 *   an evaluation at rho. */
int ofdft_local_pp_forces(ofdft_plan *plan, const void *rho, int nspecies, const double *q0, const double *dq,
                          const int *ncoef, const double *const *coef, const double *qmax, const double *vsmall,
                          const double *dqs, int nat, const double *pos_cart, const double *pos_frac,
                          const int *species, double *forces, void *stream);
int ofdft_local_pp_stress(ofdft_plan *plan, const void *rho, int nspecies, const double *q0, const double *dq,
                          const int *ncoef, const double *const *coef, const double *qmax, const double *vsmall,
                          const double *dqs, int nat, const double *pos_cart, const double *pos_frac,
                          const int *species, double energy, double *stress, void *stream);
int ofdft_ewald_forces_stress(ofdft_plan *plan, int nat, const double *pos_cart, const double *pos_frac,
                              const double *charges, double precision, double *forces, double *stress, void *stream);
/* PME forms of the same (the reference's defaults, PME=True): order-`order`
 * cardinal B-splines on the plan's grid.
 * ofdft_local_pp_forces_pme: LocalPseudo._ForcePME (pseudo/__init__.py:475-520),
 *   per species the field irfftn(conj(B) vloc_s rho(G)) (G = 0 dropped)
 *   contracted with each atom's stencil derivative as the reference does,
 *   matmul(dQ/d(frac)^T, A^-1) (rows of A = lattice vectors).
 * ofdft_local_pp_stress_pme: LocalPseudo._StressPME (pseudo/__init__.py:522-563):
 *   the exact form with each species' structure factor B rfftn(Q_s).
 * ofdft_ewald_forces_stress_pme: ewald.forces / ewald.stress with PME=True
 *   (ewald.py:510-539): real-space sums plus Forces_rec_PME (810-864) and
 *   Stress_rec_PME (915-960). */
int ofdft_local_pp_forces_pme(ofdft_plan *plan, const void *rho, int nspecies, const double *q0, const double *dq,
                              const int *ncoef, const double *const *coef, const double *qmax, const double *vsmall,
                              const double *dqs, int nat, const double *pos_cart, const double *pos_frac,
                              const int *species, int order, double *forces, void *stream);
int ofdft_local_pp_stress_pme(ofdft_plan *plan, const void *rho, int nspecies, const double *q0, const double *dq,
                              const int *ncoef, const double *const *coef, const double *qmax, const double *vsmall,
                              const double *dqs, int nat, const double *pos_cart, const double *pos_frac,
                              const int *species, int order, double energy, double *stress, void *stream);
int ofdft_ewald_forces_stress_pme(ofdft_plan *plan, int nat, const double *pos_cart, const double *pos_frac,
                                  const double *charges, double precision, int order, double *forces, double *stress,
                                  void *stream);
int ofdft_functional_stress(ofdft_plan *plan, const void *rho, const double *energies, double *out, void *stream);

/* Start densities from atomic radial functions (density/density.py).
 * ofdft_atomic_density: the real-space placement of AtomicDensity.step
 *   (:18-67) as guess_rho_atom (:209-249, kind 0) and build_pseudo_density /
 *   gen_gaussian_density (:353-406, kind 1) use it, added into rho (device,
 *   N doubles, FP64 plan).  Per species k: rcut[k], rtol[k] (kind 0: points
 *   with r <= rtol skipped when rtol > 0), scale[k] and, kind 1, sigma[k]:
 *   value scale * exp(-r^2 / 2 sigma^2) / (sqrt(2 pi) sigma)^3; kind 0: the
 *   piecewise cubic of nbreak[k] increasing breakpoints breaks[k] and
 *   4 (nbreak[k] - 1) coefficients coef[k] (c0..c3 in powers of r - x_i).
 *   Each atom covers min(int(rcut / spacing) + 1, nr / 2) points per side of
 *   the grid point below it; displacements are formed as density.py:55 does
 *   (lattice[k][j] s_j); where the box reaches a grid point twice the later
 *   (+border) value is kept, as the reference's fancy-index += does.
 *   pos_frac: nat x 3 scaled positions; species: nat indices.  Setup call
 *   (returns after the placement).
 * ofdft_clamp_normalize: rho < dtol -> dtol when clamp, then, if
 *   ncharge > 1e-6, rho *= ncharge / (sum rho dV) (density.py:236-248,
 *   300-311); FP64 plan, stream-ordered. */
int ofdft_atomic_density(ofdft_plan *plan, int kind, int nspecies, const int *nbreak, const double *const *breaks,
                         const double *const *coef, const double *rcut, const double *rtol, const double *sigma,
                         const double *scale, int nat, const double *pos_frac, const int *species, void *rho,
                         void *stream);
int ofdft_clamp_normalize(ofdft_plan *plan, void *rho, int clamp, double dtol, double ncharge, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (serving loops, harnesses that forbid hipMalloc): the same placement
 * as ofdft_atomic_density, same arguments and result, but the upload of the
 * species records, spline tables, positions and lattice goes into a static
 * device buffer of 4 MiB (one per GPU, held until the process ends) instead
 * of a per-call allocation; no hipMalloc/hipFree.  Inputs whose upload needs
 * more than 4 MiB (about 150,000 atoms) are rejected with an error.
 * Every call on one GPU shares that buffer: calls that overlap on one GPU
 * (other threads or streams) race; each call returns after its placement. */
int ofdft_atomic_density_noalloc(ofdft_plan *plan, int kind, int nspecies, const int *nbreak,
                                 const double *const *breaks, const double *const *coef, const double *rcut,
                                 const double *rtol, const double *sigma, const double *scale, int nat,
                                 const double *pos_frac, const int *species, void *rho, void *stream);

/* Preconditioner of the TN inner conjugate-gradient solve (no counterpart in
 * the reference, whose get_direction_TN, optimization.py:91-147, runs plain
 * CG).  table (host, n doubles): M^-1(q) at q = i dq, linearly interpolated
 * in |G| and held at its last value beyond the table; while set (n > 0),
 * ofdft_optimize's TN applies z = IFFT[M^-1(|G|) FFT[r]] to each CG residual
 * (preconditioned CG; the stopping tests stay on the unpreconditioned
 * residual).  n <= 0 removes it.  Setup call (copies the table). */
int ofdft_set_tn_precond(ofdft_plan *plan, int n, double dq, const double *table);
/* The same preconditioner with its table built on the device, in stream
 * order (no copy, no host table): M(q) = 2 rho0 K(q) at q = i dq, i < n, with
 * the uniform-gas response K(q) = resp[0] + resp[1] q^2 + resp[2] G(q /
 * resp[3]) + resp[5] 4 pi / q^2, G the inverse Lindhard function with
 * lambda = resp[4]; M^-1 = 0 where M is infinite. */
int ofdft_set_tn_precond_uniform(ofdft_plan *plan, int n, double dq, double rho0, const double *resp,
                                 void *stream);
/* Copy the first n values of the plan's preconditioner table to out (device,
 * stream-ordered); n at most the table's length. */
int ofdft_get_tn_precond(ofdft_plan *plan, double *out, int n, void *stream);
enum {
    OFDFT_OPT_TN = 0,
    OFDFT_OPT_CG_HS,
    OFDFT_OPT_CG_PR,
    OFDFT_OPT_CG_FR,
    OFDFT_OPT_CG_DY,
    OFDFT_OPT_CG_CD,
    OFDFT_OPT_CG_LS,
    OFDFT_OPT_GD,
    OFDFT_OPT_LBFGS
};
/* Density optimizer (Optimization.optimize_rho, optimization/optimization.py:281-520).
 * method: OFDFT_OPT_TN (get_direction_TN :91-147), the CG directions of
 * math_utils.py:395-423 (CG-HS, -PR, -FR, -DY, -CD, -LS, with the reference's
 * CG-LS denominator sum(d r0) and its field-valued CG-CD beta), OFDFT_OPT_GD
 * (-residual, :426-428) or OFDFT_OPT_LBFGS (two-loop recursion :431-455 with
 * at most maxcor + 1 pairs, :374-392, kept in the plan's pair memory that
 * ofdft_plan_reserve_lbfgs sets up; an L-BFGS call without it fails).  Line search: Moré-Thuente dcsrch on theta in
 * [0, pi] (math_utils.py:76-106).  hvp: 0 = forward difference with
 * eps = 1e-9 as the reference; 1 = analytic linearised potential (exact
 * Hessian-vector product, used in FP32).  rho: device, N reals, overwritten
 * with the optimised density; pot (device, N) receives the final potential.
 * opts (host, 10 doubles): econv, maxiter, maxfun, maxls, c1, c2, xtol, ncheck,
 * maxcor, h0 (the last two read by L-BFGS only; h0 = 0 scales the first loop's
 * result by s.y / y.y of the newest pair, as a falsy H0 does in the reference).
 * stats (host, 16 doubles): [converged, niter, energy, mu, n_eval_v, n_eval_ev,
 * seconds, ...]; converged: 0 converged, 1 line search failed, 2 maxiter
 * reached (optimization.py:440-444,509-511).  FP32 plans skip the line
 * search when |dE/dtheta| theta of the trial step is below econv / 10 (its
 * energy differences would sit at single-precision resolution) and take the
 * trial step.
 * history (host, 2*hcap doubles): energies per step in [0, hcap), seconds since
 * the call started at each step in [hcap, 2*hcap). */
int ofdft_optimize(ofdft_plan *plan, void *rho, void *pot, int method, int hvp, const double *opts,
                   double *stats, double *history, int hcap, void *stream);
/* The TN form of ofdft_optimize (same arguments and results; method TN) with
 * the iteration driven from the device: the inner-CG stopping tests
 * (optimization.py:113-145), the uphill restart, the Moré-Thuente line search
 * (the same arithmetic as ofdft_dcsrch) and the convergence test
 * (optimization.py:480-508) run in single-thread control kernels, and the
 * kernels of one superstep -- CG start, cg_steps inner-CG iterations, the
 * orthonormalisation, ls_steps line-search evaluations and the accept, each
 * skipped when its control flag is off -- are captured once as a hipGraph
 * (cached on the plan, updated in place when the lattice or fields change)
 * and launched with one read-back of a few scalars per superstep.  The
 * results equal ofdft_optimize's up to the last bits of cos/sin of the line
 * search step (device vs host libm).  stats[7] = supersteps launched,
 * stats[8], stats[9] = graph captures / in-place updates on this plan so far.
 * Needs size-specialised FFT passes on every axis and no LMGP term.  The
 * first call on a plan (and any call after its settings change) captures;
 * the capture stream is created once per plan (default priority).
 * No counterpart in the reference beyond get_direction_TN / optimize_rho. */
int ofdft_optimize_tn_graph(ofdft_plan *plan, void *rho, void *pot, int hvp, const double *opts, int cg_steps,
                            int ls_steps, double *stats, double *history, int hcap, void *stream);
/* Setup call for OFDFT_OPT_LBFGS (no separate counterpart: the reference's
 * LBFGS class, math_utils.py:374-392, keeps its pairs in Python lists): makes
 * the plan hold memory for maxcor + 1 (s, y) pairs, 2 (maxcor + 1) fields of
 * the plan precision (0 <= maxcor <= 15); grows only, freed with the plan. */
int ofdft_plan_reserve_lbfgs(ofdft_plan *plan, int maxcor);

/* Moré-Thuente line search step (MINPACK-2 dcsrch, reached through
 * scipy in math_utils.py:6-9,76-106): reverse-communication state machine
 * on the host.  task: 0 START, 1 FG, 2 CONVERGENCE, 3 WARNING, 4 ERROR.
 * state: 32 doubles owned by the caller. */
int ofdft_dcsrch(double *stp, double f, double g, double ftol, double gtol, double xtol,
                 double stpmin, double stpmax, int *task, double *state);

/* TD-OFDFT Crank-Nicolson step (td/propagator/crank_nicholson.py:147-188):
 * solves (1 + i H dt/2) psi1 = (1 - i H dt/2) psi0 with the complex-
 * symmetric CG of linear_solver.py:24-47, H psi = -1/2 lap_sigma(psi) + v psi
 * (td/hamiltonian.py:70-114, field.py:343-350).  psi0/psi1: device full
 * complex fields (nx,ny,nz); v: device real potential; inner products FP64.
 * info (host, 4 doubles): [status, iterations, final residual, atol]. */
int ofdft_cn_step(ofdft_plan *plan, const void *psi0, const void *v, double dt, double sigma,
                  double tol, int maxiter, void *psi1, double *info, void *stream);
/* The same step with the linear solver chosen as crank_nicholson.py:65-77
 * offers: solver 0 cg (as ofdft_cn_step), 1 bicg and 2 bicgstab
 * (linear_solver.py:49-110, restated), 3 GMRES(restart) (the gmres_scipy
 * choice; Arnoldi by classical Gram-Schmidt applied twice, FP64 Givens
 * rotations on the host, maxiter counting operator applications).  Solvers
 * 1-3 take a workspace of ofdft_cn_solver_workspace_bytes(plan, solver,
 * restart) bytes (device, caller-owned); psi1 must not alias psi0. */
size_t ofdft_cn_solver_workspace_bytes(const ofdft_plan *plan, int solver, int restart);
int ofdft_cn_step_solver(ofdft_plan *plan, int solver, int restart, const void *psi0, const void *v, double dt,
                         double sigma, double tol, int maxiter, void *psi1, double *info, void *work,
                         size_t work_bytes, void *stream);

/* ---------------------------------------------------------------------------
 * Learned warm start (no counterpart in the reference: it produces the start
 * density handed to Optimization.optimize_rho, optimization.py:281-520, in
 * place of the guess densities of density/density.py:199-207,258-317).
 * A periodic residual 3D CNN on the density grid: 2 input channels (planar
 * FP32 fields), 32 hidden channels (channels-last bf16, N x 32), 1 output
 * channel (FP32).  Tap t = (i*3+j)*3+k reads p + dil*(i-1, j-1, k-1) with
 * periodic wrap.  nr: grid shape; every pointer is device memory.
 * ------------------------------------------------------------------------- */
/* h = relu(conv(x0; w) + b): x0 [2][N] FP32, w [27][2][32] FP32, b [32], h [N][32] bf16.
 * Computed on MFMA with x0 and w rounded to FP16 and FP32 accumulation. */
int ofdft_cnn_conv_in(const int nr[3], int dil, const float *x0, const float *w, const float *b, void *h,
                      void *stream);
/* 32 -> 32 layer on MFMA (bf16 in, FP32 accumulation), w [27][32 out][32 in] bf16.
 * mode 0 (forward): r = relu(conv(x) + bias); out0 (bf16 [N][32]) = x + r; out1 (bf16) = r, or
 *        NULL when r is not needed (inference).
 * mode 1 (backward-data, w from ofdft_cnn_prep_weights' wb): v = conv(x) + resid (FP32);
 *        out0 (FP32 [N][32]) = v; out1 (bf16) = mask > 0 ? v : 0.  out0 must not alias x
 *        (it may alias resid). */
int ofdft_cnn_conv_mid(const int nr[3], int dil, int mode, const void *x, const void *w, const float *bias,
                       const float *resid, const void *mask, void *out0, void *out1, void *stream);
/* out = conv(h; w) + b: h [N][32] bf16, w [27][32] FP32, b [1], out [N] FP32.
 * Each point's 27 tap projections are computed on FP16 MFMA (h and w rounded to
 * FP16), rounded to FP16 and summed in FP32 in tap order.  work NULL: one pass,
 * the projections of a plane kept in LDS while the layer walks the grid.  work
 * (27 N floats of capacity): two passes through 27 FP16 tap planes kept there.
 * Both forms give the same bits. */
int ofdft_cnn_conv_out(const int nr[3], int dil, const void *h, const float *w, const float *b, float *out,
                       float *work, void *stream);
/* dh [N][32] FP32 = d(sum dout*out)/dh; gout (bf16) = mask > 0 ? dh : 0 */
int ofdft_cnn_out_backward(const int nr[3], int dil, const float *dout, const float *w, const void *mask, float *dh,
                           void *gout, void *stream);
/* Floats of the partials workspace of ofdft_cnn_wgrad for `layer` and nblk blocks. */
size_t ofdft_cnn_wgrad_partials(int layer, int nblk);
/* Weight and bias gradients, reduced over all points into dw (FP32):
 * layer 0: g [N][32] bf16, x = x0 [2][N] FP32 -> dw = [27][2][32] then db [32]
 * layer 1: g [N][32] bf16, x = layer input [N][32] bf16 (MFMA) -> dw = [27][32][32] then db [32]
 * layer 2: g = dout [N] FP32, x = h [N][32] bf16 -> dw = [27][32] then db [1] */
int ofdft_cnn_wgrad(const int nr[3], int dil, int layer, const void *g, const void *x, float *partials, int nblk,
                    float *dw, void *stream);
/* wf = bf16(w) [27][32][32]; wb[t][ci][co] = bf16(w[26-t][co][ci]) for backward-data */
int ofdft_cnn_prep_weights(const float *w, void *wf, void *wb, void *stream);
/* loss (device double) = mean (pred - target)^2; dpred = 2 (pred - target) / N */
int ofdft_cnn_loss(ofdft_plan *plan, const float *pred, const float *target, float *dpred, double *loss,
                   void *stream);
/* Spectral loss (FP32 plan): with d = pred - target and D = unnormalised R2C
 * of d, loss = (1/N^2) sum over the full spectrum of w(q) |D(q)|^2 (w given
 * on the half spectrum [nx][ny][nz/2+1]; w = 1 gives mean(d^2)); dpred =
 * d loss / d pred.  Uses the plan's workspace. */
int ofdft_cnn_loss_spectral(ofdft_plan *plan, const float *pred, const float *target, const float *w, float *dpred,
                            double *loss, void *stream);
/* One Adam step (bias-corrected, FP32 master weights), step >= 1 */
int ofdft_cnn_adam(long long n, float *p, const float *g, float *m, float *v, double lr, double b1, double b2,
                   double eps, int step, void *stream);
/* Network inputs: x0[0] = c (rho_s - rho0) / rho0, x0[1] = sv * vloc (plan precision inputs) */
int ofdft_cnn_features(ofdft_plan *plan, const void *rho_s, const void *vloc, double rho0, double c, double sv,
                       float *x0, void *stream);
/* Training target: target = log(rho / rho0) - x0[0] (rho: converged density, plan precision) */
int ofdft_cnn_target(ofdft_plan *plan, const void *rho, const float *x0, double rho0, float *target, void *stream);
/* rho = rho0 exp(x0[0] + delta) scaled to ne electrons (delta may be NULL) */
int ofdft_cnn_apply(ofdft_plan *plan, const float *x0, const float *delta, double rho0, double ne, void *rho,
                    void *stream);

/* Crank-Nicolson step for nb independent runs at once (1 <= nb <= 16; no
 * counterpart in the reference, which propagates one run per process:
 * td/propagator/crank_nicholson.py:147-188 per run).  psi0/psi1: nb stacked
 * full complex fields, v: nb stacked real potentials, info (host, 4*nb):
 * per run as ofdft_cn_step.  Each run keeps its own CG scalars and stops
 * when it converges.  work: device buffer of at least
 * ofdft_cn_batch_workspace_bytes(plan, nb) bytes. */
size_t ofdft_cn_batch_workspace_bytes(const ofdft_plan *plan, int nb);
int ofdft_cn_step_batch(ofdft_plan *plan, int nb, const void *psi0, const void *v, double dt, double sigma,
                        double tol, int maxiter, void *psi1, double *info, void *work, size_t work_bytes,
                        void *stream);

#ifdef __cplusplus
}
#endif

#endif /* OFDFT_HIP_H */
