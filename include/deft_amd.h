/*
 * deft_amd: C ABI of the MI355X White-Bear fundamental-measure functional and
 * its conjugate-gradient minimizer (one grid, resident on one GPU).
 *
 * The reference (deft, github.com/droundy/deft) reaches this path through three
 * nested seams; each entry point below names the one it replaces:
 *   - the FFT seam        fft()/ifft()                 src/Grid.cpp:470-478, src/ReciprocalGrid.cpp:52-65,
 *                                                      src/new/Vector.h:327-373
 *   - the functional seam FunctionalInterface          src/Functional.h:9-42 (integral, grad),
 *                         generated NewFunctional      src/new/NewFunctional.h:27-87
 *                                                      (energy, grad, energy_grad_and_precond)
 *   - the minimizer seam  Minimize / MinimizerInterface src/new/Minimize.h:13-199, src/Minimizer.h:16-81
 *
 * Conventions (the reference's, kept):
 *   - real fields are row-major, index (x*Ny + y)*Nz + z (src/Grid.h:37-42);
 *     half spectra (x*Ny + y)*(Nz/2+1) + kz, interleaved re/im doubles;
 *   - any grid edge from 1 to 4096 whose prime factors are at most 512, as FFTW
 *     takes any N (src/GridDescription.cpp:26-30): the edges of the old-API
 *     drivers (65 = 1 + int(6.4/0.1 + e/100), 201, the 1 x 1 x 3601 wall lines
 *     of papers/contact/figs/walls.cpp:91-93) as well as the even ones the new
 *     API asserts (src/new/Vector.h:328-330).  Edges with fused per-edge
 *     kernels (deft_fused_length) run the fused passes; the others run line
 *     passes with a runtime radix schedule;
 *   - gradients are ACCUMULATED into caller buffers (+=), and the pgrad output
 *     may be NULL (src/Functional.h:407-417, src/EffectivePotentialToDensity.cpp:62-79);
 *   - gradients are derivatives of the discrete energy sum, so they carry the dV
 *     factor (ingrad = dV, src/Functional.h:145-148);
 *   - NaN and infinity are values, not errors; minimizers stop on them
 *     (src/new/Minimize.cpp:22-30,75-111).
 * Every function returns a deft_status; deft_last_error() describes the last
 * failure on the calling thread.  Plans and minimizers are not thread-safe
 * (the reference is single-threaded too); distinct plans may be used from
 * distinct threads.
 */
#ifndef DEFT_AMD_H
#define DEFT_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  DEFT_OK = 0,
  DEFT_ERR_ARG = 1,          /* invalid argument */
  DEFT_ERR_UNSUPPORTED = 2,  /* grid edge not supported by the FFT kernels */
  DEFT_ERR_HIP = 3,          /* HIP runtime error (no device, launch failure, ...) */
  DEFT_ERR_ALLOC = 4,        /* device or pinned allocation failed */
  DEFT_ERR_STATE = 5         /* call not valid in the current state */
} deft_status;

/* FFT / convolution precision.  Pointwise Phi, log(1-n3), 1/(1-n3), all
 * reductions and the CG scalars are FP64 in both modes. */
typedef enum {
  DEFT_FFT_FP32 = 0,
  DEFT_FFT_FP64 = 1
} deft_fft_precision;

/* Hard-sphere functional of a plan: White-Bear without tensor
 * (HardSpheresWBnotensor, src/HardSpheres.cpp:137-159, the default),
 * White-Bear Mark II (WhiteBearMarkII, src/haskell/WhiteBear.hs:140-159), or
 * Mark II as the C++ HardSpheresWBm2slow writes it (src/HardSpheres.cpp:161-185:
 * Phi3's prefactor 1/(36 pi) in place of 1/(24 pi)), the form the reference's
 * published Mark II profiles (papers/contact/figs/inner-sphereWBm2-*.dat) match.
 * The first two give the Carnahan-Starling bulk that deft_bulk_chemical_potential
 * and friends compute; deft_bulk_state gives the bulk of any of them.
 * Batches (deft_batch_*) take White-Bear plans only. */
typedef enum {
  DEFT_WHITE_BEAR = 0,
  DEFT_WHITE_BEAR_MARK_II = 1,
  DEFT_WHITE_BEAR_MARK_II_SLOW = 2,
  /* TensorWhiteBear (src/haskell/WhiteBear.hs:44-51, src/HardSpheres.cpp:30-67):
   * Phi3 with the shell tensor density (six more weighted fields); same bulk as
   * White-Bear */
  DEFT_WHITE_BEAR_TENSOR = 3,
  /* Soft fundamental-measure theory (SFMTFluidVeff, src/haskell/SFMT.hs:13-133):
   * Gaussian-smeared independent weights w3, w2, w1, w0, w2v, w1v and the
   * transverse tensor weight (sixteen weighted fields), softness from
   * deft_plan_set_sfmt.  Runs on the generic line passes; no batches, no
   * structure-factor preconditioner, no contact estimators. */
  DEFT_SFMT = 4,
  /* The square-well liquid (SW_liquidVeff, src/haskell/SW_liquid.hs): White-Bear
   * plus the attraction (eps/2) int n [sum_i Xi_i * (n P_i(gSigmaA))] of a well
   * of depth eps and range lambda sigma (sigma = 2R), with the White-Bear
   * gSigmaA of the same density; parameters from deft_plan_set_sw.  Runs on
   * the generic line passes; no batches, no structure-factor preconditioner,
   * no contact estimators; Vext finite where the fluid is (the reference's
   * drivers cap it, papers/square-well-fluid/figs/new-radial-sw.cpp:142-150). */
  DEFT_SW_LIQUID = 5
} deft_functional_kind;

/* Why a minimizer stopped (mirrors the return-false branches of
 * src/new/Minimize.cpp:9-313). */
typedef enum {
  DEFT_STOP_NONE = 0,
  DEFT_STOP_MAXITER = 1,
  DEFT_STOP_NAN_ENERGY = 2,
  DEFT_STOP_INFINITE_ENERGY = 3,
  DEFT_STOP_NAN_SLOPE = 4,
  DEFT_STOP_ZERO_SLOPE = 5,
  DEFT_STOP_NEGATIVE_CURVATURE = 6,
  DEFT_STOP_ENERGY_ROSE = 7,
  DEFT_STOP_NO_CHANGE = 8,
  DEFT_STOP_CONVERGED = 9,
  /* the old-API conjugate gradient did not lower the energy (its
   * improve_energy returned false, src/ConjugateGradient.cpp:107,159) */
  DEFT_STOP_NOT_IMPROVED = 10,
  /* the precision ladder's final verdict was not confirmed by the gradient
   * at the final point (no counterpart in the reference) */
  DEFT_STOP_GRADIENT = 11
} deft_stop_reason;

/* Preconditioned gradient of deft_functional_grad_kind:
 *   DEFT_PGRAD_NEW                 grad/n = -(dF/dn)/kT, the new API's
 *                                  energy_grad_and_precond (src/haskell/NewCode.hs:229-242)
 *   DEFT_PGRAD_EFFECTIVE_POTENTIAL (dF/dn)/(Veff - kT), what the old API's
 *                                  OfEffectivePotential::grad accumulates into outpgrad
 *                                  (src/EffectivePotentialToDensity.cpp:62-71)
 *   DEFT_PGRAD_STRUCTURE_FACTOR    no counterpart in the reference:
 *                                  n^-1/2 F^-1 S(k) F n^-1/2 grad with S(k) the
 *                                  structure factor of the uniform fluid at the bulk
 *                                  density of the plan's mu (deft_plan_bulk_density),
 *                                  on points with n >= 1e-3 n_b; grad/n below that,
 *                                  zero on hard-wall and constrained points.  In the
 *                                  uniform fluid it is dV/kT times the inverse Hessian
 *                                  over Veff; with S = 1 it is DEFT_PGRAD_NEW.  Costs one
 *                                  extra forward and inverse 3-D FFT per gradient. */
typedef enum {
  DEFT_PGRAD_NEW = 1,
  DEFT_PGRAD_EFFECTIVE_POTENTIAL = 2,
  DEFT_PGRAD_STRUCTURE_FACTOR = 3
} deft_pgrad_kind;

typedef struct deft_plan_s *deft_plan;
typedef struct deft_minimize_s *deft_minimize;
typedef struct deft_cg_s *deft_cg;

/* Per-term energies of the last evaluation, like the reference's printme /
 * print_summary (src/Functional.cpp:281-316): kT*int Phi1, kT*int Phi2,
 * kT*int Phi3, ideal gas, int Vext n, -mu N, and their sum. */
typedef struct {
  double kTphi1, kTphi2, kTphi3, ideal, int_V_n, minus_mu_N, total;
  double sw; /* square-well attraction (DEFT_SW_LIQUID; 0 otherwise), included in total */
} deft_energy_terms;

typedef struct {
  int iter;                 /* Minimize::get_iteration_count (src/new/Minimize.h:77) */
  int num_energy_calcs;
  int num_grad_calcs;
  int stop_reason;          /* deft_stop_reason */
  double energy;            /* last (cached) energy */
  double delta_energy;      /* energy change of the last iteration */
  double error_estimate;    /* src/new/Minimize.cpp:263-312 */
  double step;              /* remembered line-search step */
} deft_minimize_stats;

const char *deft_last_error(void);
int deft_version(void);
/* 1 if an FFT edge of this length is supported, else 0. */
int deft_supported_length(int n);
/* 1 if the fused per-edge kernels are compiled for this edge
 * (16, 20, 21, 32, 48, 64, 96, 128, 201, 256), else 0.  A plan whose three
 * edges are all fused runs the fused passes and may join a batch; any other
 * plan runs the generic line passes. */
int deft_fused_length(int n);
/* Number of visible HIP devices (0 without a GPU). */
int deft_device_count(int *count);

/* --- plan: grid + White-Bear fluid parameters -------------------------------
 * Replaces GridDescription (src/GridDescription.cpp:6-50) plus the generated
 * WhiteBearFluidVeff constructor (src/haskell/NewCode.hs:150-178):
 *   cell  : 3x3 row-major, rows are the lattice vectors a1, a2, a3 (src/lattice.h:79-112)
 *   R, kT : hard-sphere radius and temperature; mu: chemical potential
 * F[Veff] = kT int Phi_WB + int kT n (ln(n/nQ) - 1) + int Vext n - mu int n,
 * n = exp(-Veff/kT) (src/haskell/WhiteBear.hs:29-31).  Vext defaults to 0;
 * Vext = +inf marks a hard wall (density exactly 0 there).  The plan owns a
 * device-resident Veff (the minimization variable), initialised to 0. */
int deft_plan_create(int nx, int ny, int nz, const double cell[9], double R, double kT, double mu,
                     int fft_precision, deft_plan *out);
int deft_plan_destroy(deft_plan plan);
/* Autotuner (no counterpart in the reference, whose FFTW_MEASURE planning,
 * src/GridDescription.cpp:19-25, plays the same role): times the instantiated
 * variants of the z pass (block of 512 or 1024 complex values), the y pass and
 * the x passes (tiles of 8 or 16 KiB) on this grid for both FFT precisions and
 * keeps the fastest; times the candidate radix orders of every axis for the
 * generic line passes; and, for a plan whose edges have fused kernels
 * (outside batches, not the tensor functional), times one FP32 and one FP64
 * energy+gradient evaluation with each engine -- the fused passes with their
 * compile-time radix order against the generic passes at their tuned orders
 * -- and keeps the faster.  Overwrites the plan's spectral scratch, not Veff
 * or the gradient buffers. */
int deft_plan_autotune(deft_plan plan);
/* The plan's engine (1: generic line passes, 0: fused per-edge passes) and the
 * two evaluation times of the last autotune (-1 when not timed); and an
 * explicit choice (0 needs fused kernels for every edge; plans outside
 * batches).  No counterpart in the reference. */
int deft_plan_engine(deft_plan plan, int *generic, double *fused_ms, double *generic_ms);
int deft_plan_set_engine(deft_plan plan, int generic);
/* The chosen variants for one FFT precision, and an explicit choice
 * (zblock 512 or 1024, ytile and xtile 8192 or 16384). */
int deft_plan_tuning(deft_plan plan, int fft_precision, int *zblock, int *ytile, int *xtile);
int deft_plan_set_tuning(deft_plan plan, int fft_precision, int zblock, int ytile, int xtile);
/* The White-Bear Phi pass variant for one FFT precision: 0 transforms a
 * block's lines with block-wide FFT stages, 1 gives each field to one wave
 * (fused plans of the White-Bear functional with a power-of-two z edge,
 * 16..512; it runs at zblock 512, with zblock 1024 the block-wide pass runs);
 * both return identical bits.  Set by deft_plan_autotune.  No counterpart in
 * the reference. */
int deft_plan_zphi_variant(deft_plan plan, int fft_precision, int *variant);
int deft_plan_set_zphi_variant(deft_plan plan, int fft_precision, int variant);
/* The forward x pass variant (the transform of the density and the five
 * weighted densities): 0 block-wide FFT stages, 1 one tile of columns per wave
 * (fused plans of the five-field functionals with a power-of-two x edge);
 * both compute the same values, to rounding.  Set by deft_plan_autotune.  No
 * counterpart in the reference. */
int deft_plan_xfwd_variant(deft_plan plan, int fft_precision, int *variant);
int deft_plan_set_xfwd_variant(deft_plan plan, int fft_precision, int variant);
/* Radix schedule of one axis (0 x, 1 y, 2 z) for one FFT precision, in
 * Stockham stage order: radix 8, 4, 2, 3, 5 butterflies in registers, other
 * primes as direct-DFT stages.  Plans with a non-fused edge take any valid
 * schedule (4, 8 or primes up to 512, product = the edge, at most 12 stages),
 * and deft_plan_autotune times the candidate orders of every axis; fused
 * plans report their compile-time order and refuse a change
 * (DEFT_ERR_UNSUPPORTED).  *nstages is the stage count; at most `capacity`
 * radices are written. */
int deft_plan_radix_schedule(deft_plan plan, int fft_precision, int axis, int *radices, int capacity, int *nstages);
int deft_plan_set_radix_schedule(deft_plan plan, int fft_precision, int axis, const int *radices, int nstages);
int deft_plan_set_fft_precision(deft_plan plan, int fft_precision);
int deft_plan_set_mu(deft_plan plan, double mu);
/* The bulk density whose chemical potential (ideal gas + the plan's
 * functional, deft_bulk_state) is the plan's mu. */
int deft_plan_bulk_density(deft_plan plan, double *nb);
/* One of the deft_functional_kind values, default DEFT_WHITE_BEAR; drops nothing else. */
int deft_plan_set_functional(deft_plan plan, int functional);
/* Soft FMT parameters of the generated SFMTFluidVeff's sigma(), epsilon() and
 * Xi() (src/haskell/SFMT.hs:43-56): alpha = sigma (2/(1 + sqrt(ln 2 kT/epsilon)))^(1/6)
 * with the plan's kT.  Defaults sigma = epsilon = 1, Xi = 0.1.  All three must
 * be positive and finite. */
int deft_plan_set_sfmt(deft_plan plan, double sigma, double epsilon, double xi);
/* HomogeneousSFMTFluid (src/haskell/SFMT.hs:39-41): free energy per volume
 * kT Phi(n) + kT n (ln(n/nQ) - 1) - mu n of the uniform fluid at density n and
 * its derivative d_by_dn; either output may be NULL.  Host FP64. */
int deft_sfmt_homogeneous(double sigma, double epsilon, double xi, double kT, double n, double mu,
                          double *energy, double *d_by_dn);
/* Square-well parameters of the generated SW_liquidVeff's epsilon() and
 * lambda() (src/haskell/SW_liquid.hs:13-27); R() is the plan's radius.
 * Defaults epsilon = 1, lambda = 1.3.  epsilon finite, lambda > 1. */
int deft_plan_set_sw(deft_plan plan, double epsilon, double lambda);
/* HomogeneousSW_liquid (src/haskell/SW_liquid.hs:45-47): free energy per
 * volume of the uniform liquid at density n (White-Bear + ideal gas + the
 * attraction with the bulk gSigmaA, - mu n) and d_by_dn; either output may be
 * NULL.  Host FP64. */
int deft_sw_homogeneous(double R, double epsilon, double lambda, double kT, double n, double mu, double *energy,
                        double *d_by_dn);
/* Vext (N doubles, host); kept on the device in FP64 for the FP64 passes and
 * rounded to FP32 for the FP32 passes.  +inf marks a hard wall
 * (density exactly 0); NaN marks a constrained point as the old API's
 * constrain(mask, f) (src/Functional.cpp:759-838) treats mask = 0: its density
 * exp(-Veff/kT) counts in the energy (no external potential), its gradient and
 * pgrad are zero, so its Veff stays where the caller put it. */
int deft_plan_set_vext(deft_plan plan, const double *vext);
int deft_plan_set_veff(deft_plan plan, const double *veff);
/* Veff = value at every point, on the device (the uniform start -kT ln n_b of
 * deft's drivers, e.g. papers/contact/figs/inner-sphere.cpp:197-198, without
 * a host upload). */
int deft_plan_fill_veff(deft_plan plan, double value);
int deft_plan_get_veff(deft_plan plan, double *veff);
/* n = exp(-Veff/kT), 0 on hard-wall points */
int deft_plan_get_density(deft_plan plan, double *n);
/* Device pointer of the resident Veff (N doubles) for zero-copy callers. */
int deft_plan_veff_device_ptr(deft_plan plan, double **ptr);
int deft_plan_sizes(deft_plan plan, long long *n_real, long long *n_spectral);

/* --- functional seam -------------------------------------------------------
 * FunctionalInterface::integral (src/Functional.h:131-143, src/Functional.cpp:67-69) /
 * NewFunctional::energy.  veff may be NULL to use the resident Veff; otherwise
 * it is uploaded and becomes the resident Veff. */
int deft_functional_energy(deft_plan plan, const double *veff, double *energy, deft_energy_terms *terms);
/* The generated new-API class's energy_grad_and_precond (src/haskell/NewCode.hs:229-242)
 * in the accumulate form of FunctionalInterface::grad (src/Functional.h:147-149,184-188),
 * with the OfEffectivePotential chain rule (src/EffectivePotentialToDensity.cpp:62-79):
 *   outgrad  += dF/dVeff = -(n/kT) dF/dn
 *   outpgrad += grad/n = -dF/dn / kT  (new-API preconditioner)
 * outpgrad and energy may be NULL.  Host buffers of N doubles.  Equal to
 * deft_functional_grad_kind with DEFT_PGRAD_NEW. */
int deft_functional_grad(deft_plan plan, const double *veff, double *outgrad, double *outpgrad, double *energy);
/* Same with device buffers (veff may be NULL). */
int deft_functional_grad_device(deft_plan plan, const double *veff_dev, double *outgrad_dev, double *outpgrad_dev,
                                double *energy);
/* Functional::integralgrad(kT, gd, x, g, pg) of the old API (src/Functional.h:144-149)
 * for pgrad_kind = DEFT_PGRAD_EFFECTIVE_POTENTIAL:
 *   outgrad  += -(n/kT) dF/dn,   outpgrad += (dF/dn)/(Veff - kT)
 * (src/EffectivePotentialToDensity.cpp:62-71); constrained (Vext = NaN) and
 * hard-wall points get 0 in both, as constrain(mask, f) masks them
 * (src/Functional.cpp:809-824).  DEFT_PGRAD_NEW gives deft_functional_grad. */
int deft_functional_grad_kind(deft_plan plan, const double *veff, double *outgrad, double *outpgrad, double *energy,
                              int pgrad_kind);
int deft_functional_grad_device_kind(deft_plan plan, const double *veff_dev, double *outgrad_dev,
                                     double *outpgrad_dev, double *energy, int pgrad_kind);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): deft_functional_grad_device_kind of a White-Bear fluid
 * given by its parameters (as deft_plan_create takes them, default tuning)
 * instead of a plan.  It calls no hipMalloc or hipFree: every intermediate
 * buffer comes from one library-owned device array of 36 MiB per GPU, held
 * from the library's load until the process ends, so grids are limited to
 * nx*ny*nz <= 64^3 = 262144 points (larger ones: DEFT_ERR_ARG).  Inputs and
 * outputs are device buffers of N doubles: veff_dev (required), vext_dev
 * (NULL: no external potential; +inf hard wall, NaN constrained point),
 * outgrad_dev (+=, required), outpgrad_dev (+=, may be NULL) and energy_dev
 * (one double, written, may be NULL).  The work is enqueued on `stream` (a
 * hipStream_t; NULL: the null stream) and the call returns without waiting.
 * All calls on one GPU share its array: calls that overlap on one GPU (other
 * streams or threads) race and must be serialised by the caller. */
int deft_functional_grad_noalloc(int nx, int ny, int nz, const double cell[9], double R, double kT, double mu,
                                 int fft_precision, int pgrad_kind, const double *veff_dev, const double *vext_dev,
                                 double *outgrad_dev, double *outpgrad_dev, double *energy_dev, void *stream);
/* The five weighted densities n3, n2, n2v_x, n2v_y, n2v_z of the resident Veff
 * (5*N doubles, field-major); n0 = n2/(4 pi R^2), n1 = n2/(4 pi R), n1v = n2v/(4 pi R).
 * Tensor White-Bear plans add the shell tensor T_xx, T_yy, T_zz, T_xy, T_yz, T_zx (11*N). */
int deft_weighted_densities(deft_plan plan, double *out5);
/* Contact-density estimators of the resident density, the six fields the
 * contact-paper drivers write beside it (papers/contact/figs/inner-sphere.cpp:
 * 220-233, radial_plot columns 3-8): n0, gSigmaS, YuWu_S, nA, gSigmaA, Gross
 * (src/haskell/WhiteBear.hs:68-142 gSigmaS/gSigmaA/nA/correlation_gross with the
 * plan's White-Bear or Mark II Phi, src/ContactDensity.cpp:362-400 YuWu), then
 * gSigmaS as the C++ by-hand gSigmaScpp / gSigmaSm2cpp writes it
 * (src/ContactDensity.cpp:231-355).  FP64; 7*N doubles, field-major; NaN where
 * n2 = 0 (deep inside hard walls).  Not for the tensor functional. */
int deft_contact_values(deft_plan plan, double *out7);

/* --- FFT seam --------------------------------------------------------------
 * fft(gd, f) = r2c * dV (src/Grid.cpp:470-478): in N doubles, out 2*Nk doubles.
 * ifft(gd, F) = c2r / V (src/ReciprocalGrid.cpp:52-65): in 2*Nk doubles, out N.
 * Run at the plan's FFT precision. */
int deft_fft(deft_plan plan, const double *in, double *out);
int deft_ifft(deft_plan plan, const double *in, double *out);

/* --- bulk (homogeneous) path, host-side FP64 --------------------------------
 * find_chemical_potential and pressure of ideal gas + White-Bear
 * (src/equation-of-state.cpp:164-192), homogeneous f(n) per volume. */
int deft_bulk_chemical_potential(double n, double R, double kT, double *mu);
int deft_bulk_pressure(double n, double R, double kT, double *p);
int deft_bulk_free_energy_density(double n, double R, double kT, double *f);
/* f(n), mu(n) and p(n) of ideal gas + the given deft_functional_kind
 * (find_chemical_potential of OfEffectivePotential(fhs + IdealGas()) in
 * papers/contact/figs/inner-sphere.cpp:181-182, for every fhs it passes). */
int deft_bulk_state(int functional, double n, double R, double kT, double *f, double *mu, double *p);

/* --- minimizer seam ---------------------------------------------------------
 * Minimize (src/new/Minimize.h:13-199): Polak-Ribiere CG with the quadratic
 * line search and cached energies (src/new/Minimize.cpp:9-313), acting on the
 * plan's resident Veff.  Defaults as the reference: miniter 6, precision 0,
 * relative_precision 1e-15, no preconditioning, step 0.1. */
int deft_minimize_create(deft_plan plan, deft_minimize *out);
int deft_minimize_destroy(deft_minimize m);
int deft_minimize_set_precision(deft_minimize m, double precision);
int deft_minimize_set_relative_precision(deft_minimize m, double relative_precision);
int deft_minimize_set_maxiter(deft_minimize m, int maxiter);
int deft_minimize_set_miniter(deft_minimize m, int miniter);
int deft_minimize_precondition(deft_minimize m, int use_preconditioning);
/* Which preconditioned gradient precondition(1) uses: DEFT_PGRAD_NEW (the
 * reference's grad/n, default) or DEFT_PGRAD_STRUCTURE_FACTOR (no counterpart
 * in the reference; in batches too, not with graph-captured rounds).  The minimum is the same; the
 * search directions, and so the iteration count, are not. */
int deft_minimize_set_preconditioner(deft_minimize m, int pgrad_kind);
/* Minimize::improve_energy: *more = 1 while the minimizer wants to continue. */
int deft_minimize_improve_energy(deft_minimize m, int verbose, int *more);
/* Loops improve_energy until it returns false or max_calls iterations ran. */
int deft_minimize_run(deft_minimize m, int max_calls, int *iterations);
/* Minimize::energy (cached). */
int deft_minimize_energy(deft_minimize m, double *energy);
int deft_minimize_stats_get(deft_minimize m, deft_minimize_stats *stats);
/* Drops cached energy/gradient (after the caller changed Veff or the plan). */
int deft_minimize_invalidate(deft_minimize m);

/* --- old-API minimizer seam -------------------------------------------------
 * MinimizerInterface (src/Minimizer.h:16-81) as built by the contact-paper
 * drivers, Precision(p, PreconditionedConjugateGradient(f, gd, kT, &x,
 * QuadraticLineMinimizer)) (papers/contact/figs/inner-sphere.cpp:205-210):
 *   deft_cg_create       ConjugateGradient(f, gd, kT, &x, QuadraticLineMinimizer, stepsize)
 *                        (preconditioned = 0, src/ConjugateGradient.cpp:59-108) or
 *                        PreconditionedConjugateGradient(...) (= 1, :115-160, with the
 *                        old-API pgrad (dF/dn)/(Veff - kT)); x is the plan's resident
 *                        Veff, which the minimizer does not own (src/Minimizer.h:19).
 *                        The reference's factories default stepsize to 10
 *                        (src/LineMinimizer.h:18-21).
 *   deft_cg_set_precision  wraps it in Precision(precision, m) (src/Precision.cpp:23-156);
 *                        0 removes the wrapper
 *   deft_cg_set_maxiter  wraps it in MaxIter(maxiter, m) (src/Minimizer.cpp); < 0 removes it
 *   deft_cg_improve_energy  MinimizerInterface::improve_energy: *more = 1 while the
 *                        outermost wrapper wants to go on
 *   deft_cg_energy, deft_cg_grad, deft_cg_pgrad  energy(), grad(), pgrad() of the
 *                        current point (grad/pgrad: N doubles into host buffers, written)
 *   deft_cg_minimize     minimize(f, gd, x): iteration count, step, direction and old
 *                        gradient start over (src/ConjugateGradient.cpp:34-44)
 * The reference recomputes every energy() call; here a point's energy is
 * computed once and reused until the point moves (same values, same
 * iterates; num_energy_calcs counts the evaluations actually run).  The
 * evaluations use the plan's FFT precision unless deft_cg_set_fft_precision
 * chooses one (-1: the plan's). */
int deft_cg_create(deft_plan plan, int preconditioned, double stepsize, deft_cg *out);
int deft_cg_destroy(deft_cg cg);
int deft_cg_set_precision(deft_cg cg, double precision);
int deft_cg_set_maxiter(deft_cg cg, int maxiter);
int deft_cg_set_fft_precision(deft_cg cg, int fft_precision);
int deft_cg_improve_energy(deft_cg cg, int verbose, int *more);
int deft_cg_energy(deft_cg cg, double *energy);
int deft_cg_grad(deft_cg cg, double *grad);
int deft_cg_pgrad(deft_cg cg, double *pgrad);
int deft_cg_minimize(deft_cg cg);
/* Drops the cached energy and gradient (after the caller changed Veff). */
int deft_cg_invalidate(deft_cg cg);
/* iter, evaluations, stop reason, energy, Precision's last delta E (old minus
 * new) and error estimate, remembered step */
int deft_cg_stats_get(deft_cg cg, deft_minimize_stats *stats);

/* --- mixed precision (no counterpart in the reference, which is FP64 only) ---
 * With switch_relative_precision > 0 the minimizer climbs a precision ladder:
 *   phase 0: energies and gradients with FP32 FFTs, until the error estimate
 *            falls below switch_relative_precision*|E| or the line search stalls;
 *   phase 1: line-search energies with FP64 FFTs, gradients FP32 (skipped when
 *            polish_grad_precision is DEFT_FFT_FP64);
 *   phase 2: everything FP64, to the minimizer's own precision, running at
 *            least 3 iterations of its own before it may report convergence.
 * Each phase hands over when the reference algorithm would stop; only phase 2
 * ends the minimization.  The convergence history carries over; the CG
 * direction restarts when the gradient precision changes.  Switch while FP32
 * energy differences are still well above FP32 round-off (about 1e-7 |E|):
 * 1e-6 is a good value.  The plan's precision setting is not changed.
 * 0 (default) disables the ladder. */
int deft_minimize_set_mixed_precision(deft_minimize m, double switch_relative_precision, int polish_grad_precision);
/* Fused line search (no counterpart in the reference): the trial point s2 of
 * the quadratic line step (src/new/Minimize.cpp:160-175) is evaluated for
 * energy and gradient at once when both use the same FFT precision, so an
 * accepted s2 costs no extra evaluation at the next iteration's start; it is
 * used while the previous quadratic step kept its s2 (a rejected s2 wastes the
 * gradient).  The energies and the iterates are those of the reference
 * algorithm; only the evaluation counts change.  mode 1 on, 0 off, -1 (default) on exactly when
 * the mixed-precision ladder is on. */
int deft_minimize_set_fused_line_search(deft_minimize m, int mode);
/* Gradient confirmation of the ladder's final verdict (no counterpart in the
 * reference, whose estimate trusts the energy changes alone): with
 * preconditioning on, a stop that phase 2 would report as converged is
 * accepted only if (kT/dV) pg.g / 2 -- the Newton estimate, from the gradient
 * at the final point, of the energy still to be gained -- is at most factor x
 * max(precision, relative_precision |E|); otherwise the CG direction restarts
 * from steepest descent (at most 4 times, then the stop is
 * DEFT_STOP_GRADIENT).  factor 0 disables it; the default is set when the
 * ladder is switched on.  Plain (non-ladder) minimizations never use it. */
int deft_minimize_set_gradient_check(deft_minimize m, double factor);
/* Newton estimate of the last confirmation (-1 if none ran) and the number of
 * restarts it caused. */
int deft_minimize_gradient_check_get(deft_minimize m, double *last_grad_error, int *restarts);
/* Iterations run in phase 0, -1 if the ladder was not used or not left. */
int deft_minimize_fp32_iterations(deft_minimize m, int *iterations);
/* Iterations of phase 0 and phase 1 and the current phase (0, 1 or 2). */
int deft_minimize_phase_iterations(deft_minimize m, int *fp32_iterations, int *mixed_iterations, int *phase);

/* --- batches (no counterpart in the reference, which runs one minimization
 * per process, e.g. papers/contact/figs/inner-sphere.cpp:177-211) -----------
 * A batch joins plans of any grid shapes (uneven edges such as {32, 48, 64,
 * 96} per axis).  deft_batch_minimize runs one minimizer per plan, each with
 * the unchanged algorithm above (and its mixed-precision ladder if set), in a
 * fiber (user-level context) of the calling thread; every evaluation and
 * every CG direction update they ask for joins a lockstep round, launched as
 * one kernel per pass and group of grids with equal axis length, with one
 * device wait per round.  deft_batch_set_graphs(batch, 1) makes a round one
 * hipGraph per precision instead, captured on the first round: every pass
 * over every grid of the batch, blocks of grids without a matching request
 * returning at once; later rounds copy the argument table and replay it.
 * Each grid gets bit for bit the energies and iterates of its own unbatched
 * run.  While in a batch a plan's work runs on the batch's stream, and the
 * plan cannot be destroyed. */
typedef struct deft_batch_s *deft_batch;
int deft_batch_create(deft_plan *plans, int n, deft_batch *out);
/* Runs every minimizer to its own stop; minimizers[i] must belong to a plan of
 * the batch, at most one per plan. */
int deft_batch_minimize(deft_batch batch, deft_minimize *minimizers, int n);
/* 1: rounds replay captured graphs; 0 (default): each round launches only the
 * requested passes, grouped afresh (faster on the 64-grid batch: the graphs'
 * idle blocks cost more than the launches they save).  Not while minimizing. */
int deft_batch_set_graphs(deft_batch batch, int on);
/* Rounds run, kernel launches of the rounds, evaluations served. */
int deft_batch_stats(deft_batch batch, long long *rounds, long long *launches, long long *requests);
/* Returns the plans to their own streams. */
int deft_batch_destroy(deft_batch batch);

/* --- profiling (the reference only has ad-hoc clock() timers,
 * src/new/Minimize.h:95-135) ------------------------------------------------
 * With profiling on, every kernel launch of the plan is bracketed by HIP
 * events on the plan's stream; per kernel family the library accumulates the
 * launch count, the summed event time and the algorithmic bytes (compulsory
 * reads + writes of that pass). */
int deft_plan_set_profiling(deft_plan plan, int on);
int deft_plan_profile_reset(deft_plan plan);
int deft_kernel_count(void);
int deft_plan_profile_get(deft_plan plan, int kernel, const char **name, long long *launches, double *total_ms,
                          double *total_bytes);
/* Attainable HBM bandwidth on the current device: a streaming copy of `bytes`
 * (16-byte vector loads and stores, four in flight per thread) timed with HIP
 * events over `reps` launches after one warm-up, best of 2-32 blocks per CU;
 * GB/s = 2 bytes / mean time.  A yardstick for the per-kernel figures above
 * (no reference counterpart). */
int deft_copy_bandwidth(long long bytes, int reps, double *gbps);

#ifdef __cplusplus
}
#endif

#endif /* DEFT_AMD_H */
