/*
 * specfab_gpu.h -- C-ABI of libspecfab_gpu, the MI355X (gfx950) implementation
 * of specfab's batched fabric-evolution and homogenization path.
 *
 * Every entry point returns an int status (SFG_OK = 0, negative on error; the
 * message is available from sfg_last_error()) instead of the reference's
 * Fortran `stop`.  The library keeps no global truncation: L is an argument of
 * every call, and a context (sfg_ctx) owns the device copies of the coupling
 * tables, so several truncations can be used from one process.
 *
 * Conventions
 *   - nlm: complex double, interleaved (re, im), specfab coefficient order
 *     (l = 0, 2, ..., L; m = -l..l), length nlm_len(L) = (L+1)(L+2)/2
 *     (reference: src/header.f90:23-33).  Batched nlm arrays are (P, n),
 *     parcel-major, i.e. the layout of specfabpy's *_arr arguments.
 *   - 3x3 tensors are row-major: A[3*i + j] = A_ij.  Batched tensors are
 *     (P, 3, 3).  (Fortran callers pass transpose(A), as the specfab_gpu module does.)
 *   - n x n operators are row-major: M[i*n + j], so d nlm_i/dt = M_ij nlm_j.
 *   - Pointers named *_dev are device (HBM) pointers; all others are host.
 *   - `stream` is a hipStream_t (NULL = HIP's default stream).  Batched calls
 *     are asynchronous on that stream; single-parcel calls use the context's
 *     own stream and synchronise before returning.
 *   - The device state of a batch is the real ("reality-reduced") SoA array
 *     x[a * P + p], a = 0..n-1: per even l the real DOFs
 *     [Re n_l0, Re n_l1, Im n_l1, ..., Re n_ll, Im n_ll] of a real ODF
 *     (n_l^-m = (-1)^m conj(n_l^m)).  sfg_nlm_to_state / sfg_state_to_nlm
 *     convert.
 */
#ifndef SPECFAB_GPU_H
#define SPECFAB_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFG_OK 0
#define SFG_ERR_ARG (-1)
#define SFG_ERR_HIP (-2)
#define SFG_ERR_L (-3)
#define SFG_ERR_NOMEM (-4)

/* process bitmask for sfg_evolve_desc.procs */
#define SFG_LROT 1
#define SFG_DDRX 2
#define SFG_CDRX 4
#define SFG_REG 8

/* time schemes */
#define SFG_EULER 0
#define SFG_RK4 1

/* precisions */
#define SFG_FP64 0
#define SFG_FP32C 1 /* FP32 stepping with compensated (two-sum) state update */

/* per-parcel status bits written by the batched routines */
#define SFG_ST_NAN 1          /* non-finite state or output */
#define SFG_ST_TAYLOR_REG 2   /* Taylor solve needed the regularized normal equations */
#define SFG_ST_TAYLOR_FAIL 4  /* ... and that failed too (reference: stop) */
#define SFG_ST_REUSS_SING 8   /* singular 9x9 Reuss matrix (reference: stop) */
#define SFG_ST_ENVELOPE 16    /* outside the surrogate's training envelope */
#define SFG_ST_AUDIT 32       /* drawn for the surrogate's random audit */
#define SFG_ST_ERRBOUND 64    /* the surrogate's Eij error bound exceeded the tolerance */

typedef struct sfg_ctx sfg_ctx;

/* Create / destroy a context on HIP device `device` (replaces initspecfab(L),
 * src/specfab.f90:37-53; the reference's init(L) of specfabpy,
 * src/specfabpy.f90:150-160). */
int sfg_ctx_create(int device, sfg_ctx** ctx);
int sfg_ctx_destroy(sfg_ctx* ctx);
const char* sfg_last_error(void);

/* Autotuner, no counterpart in the reference: times every generated
 * register-resident stepper variant of (L, precision) on a synthetic batch of
 * P parcels (launches of `steps` RK4 steps, LROT+REG+DDRX: the launch length
 * the caller will use, since a launch's prologue is paid once) and keeps the
 * fastest for this context.  *best = chosen variant (-1: no generated stepper
 * for L); ms[v] = time per step with variant v (first ms_len variants).
 * Environment variable SFG_VARIANT=k forces variant k. */
int sfg_autotune(sfg_ctx* ctx, int L, int precision, int64_t P, int steps, void* stream, int* best, double* ms,
                 int ms_len);

/* Profiling hook, no counterpart in the reference: while enabled, every
 * sfg_evolve brackets its stepping kernel with HIP events on the launch
 * stream; sfg_ctx_read_timing waits for them, returns the summed kernel time
 * (ms) and the number of launches, and clears the record. */
int sfg_ctx_set_timing(sfg_ctx* ctx, int enable);
int sfg_ctx_read_timing(sfg_ctx* ctx, double* ms_total, int* count);

/* nlm_len (src/header.f90:23) and the (l,m) list lm(2, n) (src/header.f90:30):
 * lm[2*j] = l, lm[2*j+1] = m. */
int sfg_nlm_len(int L);
int sfg_lm(int L, int* lm);

/* ---- single-parcel dense operators (host buffers) ----
 * Drop-ins for the reference operator builders; M is complex (n x n) or real
 * (n x n) row-major.  Computed on the GPU from the coupling tables. */
/* M_LROT(eps, omg, iota, zeta): src/dynamics.f90:52-97, specfabpy.f90:172-180 */
int sfg_M_LROT(sfg_ctx* ctx, int L, const double eps[9], const double omg[9], double iota, double zeta,
               double* M);
/* M_DDRX(nlm, tau): src/dynamics.f90:251-275, specfabpy.f90:182-190 */
int sfg_M_DDRX(sfg_ctx* ctx, int L, const double* nlm, const double tau[9], double* M);
/* M_DDRX_src(tau): src/dynamics.f90:277-298, specfabpy.f90:192-199 */
int sfg_M_DDRX_src(sfg_ctx* ctx, int L, const double tau[9], double* M);
/* M_CDRX(): src/dynamics.f90:474-492 (real n x n) */
int sfg_M_CDRX(sfg_ctx* ctx, int L, double* M);
/* M_REG(D): src/dynamics.f90:494-518 (real n x n) */
int sfg_M_REG(sfg_ctx* ctx, int L, const double D[9], double* M);

/* ---- batched operators (device buffers), the Eulerian per-node assembly of
 * src/specfabpy/fenics/CPO.py:200-202: M_dev is (P, n, n) complex. ---- */
int sfg_M_LROT_batch(sfg_ctx* ctx, int L, int64_t P, const double* eps_dev, const double* omg_dev, double iota,
                     double zeta, double* M_dev, void* stream);
int sfg_M_DDRX_src_batch(sfg_ctx* ctx, int L, int64_t P, const double* tau_dev, double* M_dev, void* stream);
/* M_REG(D) of every node (src/dynamics.f90:494-518; assembled per node at
 * src/specfabpy/fenics/CPO.py:202): D_dev (P,3,3) -> M_dev (P, n, n) real */
int sfg_M_REG_batch(sfg_ctx* ctx, int L, int64_t P, const double* D_dev, double* M_dev, void* stream);

/* ---- state conversion (device buffers) ---- */
int sfg_nlm_to_state(sfg_ctx* ctx, int L, int64_t P, const double* nlm_dev, double* x_dev, void* stream);
int sfg_state_to_nlm(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, double* nlm_dev, void* stream);

/* ---- reduced real form (src/reducedform.f90), the 2-D x-z Eulerian export
 * of src/specfabpy/fenics/CPO.py:200-202.  rnlm = (n_0^0, n_2^0, n_2^1,
 * n_2^2, n_4^0, ...), the m >= 0 coefficients, (L+2)^2/4 complex values; a
 * real ODF has n_l^{-m} = (-1)^m conj(n_l^m).  Even L, 0 <= L <= 30. ---- */
/* rnlm_len (src/specfabpy.f90:1039-1044); < 0 = error code */
int sfg_rnlm_len(int L);
/* reduce_M (src/reducedform.f90:76-120, bound at src/specfabpy.f90:1066-1074):
 * complex M (n, n) row-major -> real Mrr, Mri, Mir, Mii (rn, rn) row-major with
 * d(rnlm)/dt = Mrr Re r + Mri Im r + i (Mir Re r + Mii Im r) */
int sfg_reduce_M(sfg_ctx* ctx, int L, const double* M, double* Mrr, double* Mri, double* Mir, double* Mii);
/* nlm_to_rnlm (src/reducedform.f90:176-183, src/specfabpy.f90:1046-1054) */
int sfg_nlm_to_rnlm(sfg_ctx* ctx, int L, const double* nlm, double* rnlm);
/* rnlm_to_nlm (src/reducedform.f90:160-174, src/specfabpy.f90:1056-1064) */
int sfg_rnlm_to_nlm(sfg_ctx* ctx, int L, const double* rnlm, double* nlm);
/* batched over P nodes (device buffers): M_dev (P, n, n) complex as written
 * by sfg_M_LROT_batch / sfg_M_DDRX_src_batch -> Mrr..Mii_dev (P, rn, rn);
 * nlm_dev (P, n) complex <-> rnlm_dev (P, rn) complex */
int sfg_reduce_M_batch(sfg_ctx* ctx, int L, int64_t P, const double* M_dev, double* Mrr_dev, double* Mri_dev,
                       double* Mir_dev, double* Mii_dev, void* stream);
/* reduce_M(M_LROT(eps, omg, iota, zeta)) and reduce_M(M_DDRX_src(tau)) of
 * every node in one kernel, the full (n, n) matrices never formed: the same
 * values as sfg_M_LROT_batch / sfg_M_DDRX_src_batch followed by
 * sfg_reduce_M_batch (the Eulerian assembly of src/specfabpy/fenics/CPO.py:
 * 200-211); Mrr..Mii_dev (P, rn, rn). */
int sfg_M_LROT_reduced_batch(sfg_ctx* ctx, int L, int64_t P, const double* eps_dev, const double* omg_dev,
                             double iota, double zeta, double* Mrr_dev, double* Mri_dev, double* Mir_dev,
                             double* Mii_dev, void* stream);
int sfg_M_DDRX_src_reduced_batch(sfg_ctx* ctx, int L, int64_t P, const double* tau_dev, double* Mrr_dev,
                                 double* Mri_dev, double* Mir_dev, double* Mii_dev, void* stream);
int sfg_nlm_to_rnlm_batch(sfg_ctx* ctx, int L, int64_t P, const double* nlm_dev, double* rnlm_dev, void* stream);
int sfg_rnlm_to_nlm_batch(sfg_ctx* ctx, int L, int64_t P, const double* rnlm_dev, double* nlm_dev, void* stream);

/* ---- batched evolution ----
 * Steps P parcels of truncation L in place.  Per parcel: velocity gradient
 * ugrad (D = sym, W = skew), optional stress tau (NULL: tau := D, the coaxial
 * choice of src/specfabpy/integrator.py:38-39), DDRX rate Gamma0, CDRX rate
 * Lambda.  The operator per stage is M_LROT(D,W,iota,zeta) + M_REG(D) +
 * Lambda M_CDRX + Gamma0 M_DDRX(x, tau), i.e. the sum formed by
 * src/specfabpy/integrator.py:73-77, with Euler (the reference's scheme) or
 * classic RK4 (M_DDRX rebuilt from each stage's state).  steps_dev (optional,
 * int32 per parcel) overrides nsteps for uneven batches. */
typedef struct sfg_evolve_desc {
    int L;
    int64_t P;
    int procs;           /* SFG_LROT | SFG_DDRX | SFG_CDRX | SFG_REG */
    int scheme;          /* SFG_EULER or SFG_RK4 */
    int precision;       /* SFG_FP64 or SFG_FP32C */
    int nsteps;          /* uniform step count (ignored where steps_dev given) */
    double dt;
    double iota, zeta;   /* lattice-rotation parameters (src/dynamics.f90:52-73) */
    const double* ugrad_dev;  /* (P,3,3) */
    const double* tau_dev;    /* (P,3,3) or NULL */
    const double* gam0_dev;   /* (P) or NULL (0) */
    const double* lam_dev;    /* (P) or NULL (0) */
    int* steps_dev;           /* (P) or NULL; sfg_evolve only reads it, plans
                                 decrement it (remaining steps) */
    double* x_dev;            /* state, real SoA (n, P), FP64 */
    int32_t* status_dev;      /* (P) or NULL */
    /* regularization law (SFG_REG): reg_expo <= 0 (zero-initialised
     * descriptors) keeps the calibrated M_REG of L (src/include/regcalib.f90);
     * reg_expo > 0 uses -reg_nu |D| (l(l+1)/(L(L+1)))^reg_expo instead, the
     * custom law of src/specfabpy/integrator.py:85-97 (any L) */
    double reg_nu, reg_expo;
} sfg_evolve_desc;
int sfg_evolve(sfg_ctx* ctx, const sfg_evolve_desc* desc, void* stream);

/* sfg_evolve without allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, kernels
 * and results as sfg_evolve; the workspace (per-parcel launch constants and
 * the stage buffers of the chosen stepper) is a static device array of
 * SFG_NOALLOC_BYTES that each GPU the library runs on holds until the process
 * ends.  Limits: the batch must fit that workspace: 45 doubles of launch
 * constants per parcel plus the chosen stepper's stage buffers (n doubles
 * each), at L = 12 about 370,000 parcels with the FP64 register stepper
 * (no stage buffers) and 73,000 with variants that keep two stage buffers;
 * the call fails with SFG_ERR_ARG and a message otherwise.  The coupling
 * tables of desc->L must already
 * be on the device (any earlier call at that L puts them there, e.g.
 * sfg_evolve with P = 0); the call fails rather than upload them.  The
 * array's address is taken when a context is created on the device; the
 * first launch of each kernel in a process still loads its code object (the
 * HIP runtime's one-time work, as for every kernel), so make one call before
 * capturing or timing.  Every call
 * on one GPU uses the same array: calls that overlap on one GPU (from
 * several streams or threads) race.  Nothing in the array needs to start at
 * zero (every launch writes what it reads). */
#define SFG_NOALLOC_BYTES ((size_t)128 << 20)
int sfg_evolve_noalloc(sfg_ctx* ctx, const sfg_evolve_desc* desc, void* stream);

/* ---- moments and homogenization on batched states (device buffers) ---- */
/* a2 (P,3,3), src/moments.f90:37-44 */
int sfg_a2_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, double* a2_dev, void* stream);
/* a4 Mandel (P,6,6), src/moments.f90:164-178 */
int sfg_a4mandel_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, double* a4v_dev, void* stream);

/* Enhancement factors Eij (P,6) of transversely isotropic grains with a
 * linear Sachs/Taylor mix, Eij_tranisotropic (src/enhancementfactors.f90:23-69,
 * specfabpy.f90:390-400 and the *_arr loop :474-486), for P states x_dev.
 * frame: 0 = the a2 eigenframe of each parcel (eig, src/frames.f90:14-22;
 * eigenvalues -> lami_dev (P,3), eigenvectors -> ei_dev (P,3,3) rows e1,e2,e3,
 * largest eigenvalue first), 1 = caller frames frames_dev (P,3,3),
 * 2 = Cartesian (x, y, z).  Eij_grain = (Ecc, Eca), alpha = Taylor weight,
 * n_grain = 1, -3 or 3 (n' = 3 Sachs, src/homogenizations.f90:93-103: needs
 * L >= 8; its a6 / a8 terms use the exact sixth and eighth moments, the
 * reference's REAL(4)-literal bodies agree with them to ~1e-7; the Taylor
 * part keeps the n' = 1 form with n' = 3 parameters, as the reference's).
 * status_dev bits: SFG_ST_TAYLOR_REG / _FAIL / _NAN. */
int sfg_Eij_tranisotropic_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, int frame,
                                const double* frames_dev, const double Eij_grain[2], double alpha, int n_grain,
                                double* Eij_dev, double* lami_dev, double* ei_dev, int32_t* status_dev, void* stream);

/* Evw_tranisotropic(nlm, v, w, tau, Eij_grain, alpha, n_grain)
 * (src/enhancementfactors.f90:47-69; specfabpy :379-388): the enhancement
 * factor of one stress state tau (symmetric, row-major) along v (x) w,
 * (1-alpha) Sachs + alpha Taylor, normalised by the isotropic fabric.
 * Batched: the nstates states states_dev (S, 15) = v(3), w(3), tau(9) for
 * every parcel, Evw_dev (P, S); status_dev bits as for Eij. */
int sfg_Evw_tranisotropic(sfg_ctx* ctx, int L, const double* nlm, const double v[3], const double w[3],
                          const double tau[9], const double Eij_grain[2], double alpha, int n_grain, double* Evw);
int sfg_Evw_tranisotropic_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, int nstates,
                                const double* states_dev, const double Eij_grain[2], double alpha, int n_grain,
                                double* Evw_dev, int32_t* status_dev, void* stream);

/* a6(nlm), and a8 of ai(nlm) (src/moments.f90:57-66, 220-236; specfabpy
 * :601-618): the sixth and eighth moments <x^a y^b z^c> of the ODF as its
 * 28 / 45 distinct entries, monomials in the order a = d..0, b = d - a..0
 * (c = d - a - b); exact (the reference's bodies carry REAL(4) literal
 * coefficients, ~3e-8 apart).  L >= 6; rows above l = 8 do not enter.
 * Batched: m6_dev (P, 28), m8_dev (P, 45). */
int sfg_moments68(sfg_ctx* ctx, int L, const double* nlm, double m6[28], double m8[45]);
int sfg_moments68_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, double* m6_dev, double* m8_dev,
                        void* stream);

/* single-parcel eig(nlm) -> ei (3x3 rows), lami (3): src/frames.f90:14-22 */
int sfg_eig(sfg_ctx* ctx, int L, const double* nlm, double ei[9], double lami[3]);
/* eig3(M, e1, e2, e3, lami) (src/frames.f90:62-80, specfabpy :346-355): the
 * eigenpairs of a symmetric 3x3 M (row-major), ei rows e1, e2, e3, largest
 * eigenvalue first; batched M_dev (P, 9) -> ei_dev (P, 9), lami_dev (P, 3). */
int sfg_eig3(sfg_ctx* ctx, const double M[9], double ei[9], double lami[3]);
int sfg_eig3_batch(sfg_ctx* ctx, int64_t P, const double* M_dev, double* ei_dev, double* lami_dev, void* stream);
/* single-parcel Eij_tranisotropic(nlm, e1,e2,e3, Eij_grain, alpha, n_grain) */
int sfg_Eij_tranisotropic(sfg_ctx* ctx, int L, const double* nlm, const double e1[3], const double e2[3],
                          const double e3[3], const double Eij_grain[2], double alpha, int n_grain, double Eij[6]);

/* single-parcel moments (host): a2(nlm) 3x3 (src/moments.f90:37-44) */
int sfg_a2(sfg_ctx* ctx, int L, const double* nlm, double a2[9]);
/* <D> = ev_D2(nlm, tau), src/dynamics.f90:402-422 */
int sfg_ev_D2(sfg_ctx* ctx, int L, const double* nlm, const double tau[9], double* D);
/* <D> = ev_D(nlm, tau, n_RSS) (src/dynamics.f90:385-400; specfabpy
 * deformability_tranisotropic(_arr), src/specfabpy.f90:557-578): the average
 * basal resolved shear stress to the power n_RSS = 2 (ev_D2) or 4 (ev_D4,
 * src/dynamics.f90:424-449, from a4 and the a6 / a8 moments; L >= 4, the rows
 * a state lacks count as zero), normalised by its isotropic value.  Single:
 * host nlm, tau.  Batched: states x_dev (n, P), per-parcel tau_dev (P, 3, 3)
 * -> D_dev (P). */
int sfg_ev_D(sfg_ctx* ctx, int L, const double* nlm, const double tau[9], int n_RSS, double* D);
int sfg_ev_D_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, const double* tau_dev, int n_RSS,
                   double* D_dev, void* stream);
/* E_CAFFE(nlm, eps, Emin, Emax, n_grain) (src/enhancementfactors.f90:302-332;
 * specfabpy E_CAFFE(_arr), src/specfabpy.f90:461-470,543-555): the CAFFE
 * enhancement factor of the deformability <D> of the strain rate eps,
 * n_RSS = 4 for n_grain = 3 and 2 otherwise.  Batched: eps_dev (P, 3, 3)
 * -> E_dev (P). */
int sfg_E_CAFFE(sfg_ctx* ctx, int L, const double* nlm, const double eps[9], double Emin, double Emax, int n_grain,
                double* E);
int sfg_E_CAFFE_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, const double* eps_dev, double Emin,
                      double Emax, int n_grain, double* E_dev, void* stream);

/* ---- full fourth-order structure tensor and wave velocities ---- */

/* a4(nlm) (src/moments.f90:46-55, body src/include/ev_c4__body.f90 with its
 * REAL(4) literals): 3x3x3x3 row-major, a4[27i+9j+3k+l] = a4(i+1,j+1,k+1,l+1);
 * entry (3,2,1,2) is the reference body's a4(1,2,3,3).  Batched: a4_dev is
 * (P, 81). */
int sfg_a4(sfg_ctx* ctx, int L, const double* nlm, double a4[81]);
int sfg_a4_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, double* a4_dev, void* stream);

/* calibrated regularization of M_REG at L (src/include/regcalib.f90): expo, nu */
int sfg_reg_calibration(int L, double* expo, double* nu);

/* ---- initial states, DDRX rate factor, power-spectrum bounds ---- */
/* a2_to_nlm (src/moments.f90:68-74): nlm[12] = the 6 complex coefficients
 * l <= 2 in specfab order (the reference's REAL(4) literals reproduced) */
int sfg_a2_to_nlm(sfg_ctx* ctx, const double a2[9], double* nlm);
/* a4_to_nlm (src/moments.f90:76-84): nlm[30] = the 15 complex coefficients l <= 4 */
int sfg_a4_to_nlm(sfg_ctx* ctx, const double a4[81], double* nlm);
/* batched: a2 (P,3,3) resp. a4 (P,3,3,3,3) -> states x_dev (n, P) at
 * truncation L (rows above l = 2 resp. l = 4 zeroed; a4 needs L >= 4) */
int sfg_a2_to_state_batch(sfg_ctx* ctx, int L, int64_t P, const double* a2_dev, double* x_dev, void* stream);
int sfg_a4_to_state_batch(sfg_ctx* ctx, int L, int64_t P, const double* a4_dev, double* x_dev, void* stream);
/* Gamma0(eps, T, A, Q) = A sqrt(tr(eps.eps)/2) exp(-Q/(R T)) (src/dynamics.f90:451-472) */
int sfg_Gamma0(sfg_ctx* ctx, const double eps[9], double T, double A, double Q, double* G);
/* batched: eps_dev (P,3,3), temperatures T_dev (P) or the constant T (T_dev NULL) -> G_dev (P) */
int sfg_Gamma0_batch(sfg_ctx* ctx, int64_t P, const double* eps_dev, const double* T_dev, double T, double A,
                     double Q, double* G_dev, void* stream);
/* apply_bounds (src/dynamics.f90:530-557): rescale the l = 2 and l = 4
 * blocks whose power S(l) exceeds S(0); single parcel (nlm in, out: n
 * complex) and in place on batched states */
int sfg_apply_bounds(sfg_ctx* ctx, int L, const double* nlm, double* out);
int sfg_apply_bounds_batch(sfg_ctx* ctx, int L, int64_t P, double* x_dev, void* stream);

/* ---- rotation, a4 eigentensors, state validity ---- */
/* rotate_nlm(nlm, theta, phi) (src/rotation.f90:29-55, Wigner-D matrices of
 * src/include/Dlmn.f90): n'_l^m = sum_n exp(-i n phi) d^l_mn(theta) n_l^n,
 * L <= 12 (the reference's Lmax_wignerD).  Single parcel: nlm, out (n
 * complex, host).  Batched: per-parcel angles theta_dev, phi_dev (P) on the
 * states x_dev (n, P) -> out_dev (n, P), which may be x_dev. */
int sfg_rotate_nlm(sfg_ctx* ctx, int L, const double* nlm, double theta, double phi, double* out);
int sfg_rotate_nlm_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, const double* theta_dev,
                         const double* phi_dev, double* out_dev, void* stream);
/* a4_eigentensors(nlm, Q1..Q6, eigvals6) (src/frames.f90:131-152): the
 * eigenvalues (ascending) and eigentensors Q_i = vec_to_mat(v_i) of the Mandel
 * matrix of the full a4.  Single: Q[6][3][3] row-major per tensor, ev[6].
 * Batched: Q_dev (P, 6, 3, 3), ev_dev (P, 6) (either may be NULL).  L >= 4. */
int sfg_a4_eigentensors(sfg_ctx* ctx, int L, const double* nlm, double Q[54], double ev[6]);
int sfg_a4_eigentensors_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, double* Q_dev, double* ev_dev,
                              void* stream);
/* nlm_isvalid(nhat20, nhat40) (src/idealstate.f90:50-78): for K states
 * n_0^0 = 1, n_2^0 = nhat20[k], n_4^0 = nhat40[k] (complex, host arrays),
 * isvalid[k] = 1 if the a2 and a4 eigenvalues all lie in (0, 1), else 0.
 * Batched on general states: valid_dev (P) for x_dev (n, P), L >= 4. */
int sfg_nlm_isvalid(sfg_ctx* ctx, int64_t K, const double* nhat20, const double* nhat40, int32_t* isvalid);
int sfg_state_isvalid_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, int32_t* valid_dev, void* stream);

/* ---- discrete grains ---- */
/* ri_LROT(ri0, dt, Nt, D, W, iota) (src/dynamics.f90:112-137; specfabpy
 * src/specfabpy.f90:309-325): Nt - 1 Euler steps of N grain axes under
 * r' = (W + iota (r r^T D - D r r^T)) r, each step renormalised to |r| = 1.
 * ri0 (N, 3), D and W (Nt, 3, 3) row-major per step (step Nt - 1 unused),
 * ri (Nt, N, 3) with ri[0] = ri0 -- the reference's ri(Nt, N, 3) in C
 * order.  Host arrays, or device arrays on the stream (_batch); ri_dev must
 * not overlap D_dev or W_dev. */
int sfg_ri_LROT(sfg_ctx* ctx, int64_t N, int Nt, const double* ri0, double dt, const double* D, const double* W,
                double iota, double* ri);
int sfg_ri_LROT_batch(sfg_ctx* ctx, int64_t N, int Nt, const double* ri0_dev, double dt, const double* D_dev,
                      const double* W_dev, double iota, double* ri_dev, void* stream);
/* ri_to_nlm(ri, wi, L) (src/moments.f90:98-131): the harmonic expansion of a
 * weighted ensemble of N grain axes, from its a2 (L = 2), a4 (L = 4) or a6
 * (L >= 6; rows l > 6 zero) moment.  n00 = 1 / (2 sqrt(pi)) whatever the
 * weights.  Single: ri (N, 3), wi (N) host -> nlm (n complex, host).
 * Batched: E ensembles ri_dev (E, N, 3) sharing the weights wi_dev (N) ->
 * real states x_dev (n, E).  L >= 2. */
int sfg_ri_to_nlm(sfg_ctx* ctx, int L, int64_t N, const double* ri, const double* wi, double* nlm);
int sfg_ri_to_state_batch(sfg_ctx* ctx, int L, int64_t E, int64_t N, const double* ri_dev, const double* wi_dev,
                          double* x_dev, void* stream);

/* ---- orthotropic grains (olivine) ----
 * Three distributions of the slip-system axes q1 = b, q2 = n, q3 = v per
 * parcel: states x1_dev, x2_dev, x3_dev (n, P); x3 NULL or a parcel's
 * q3 n00 <= 1e-8 derives v = b x n (ai_orthotropic, src/moments.f90:357-384,
 * with the joint moments of the product density under the sin^2 kernel of
 * src/moments.f90:242-310).  Single-parcel forms take host nlm vectors
 * (nlm3 NULL or zero: derived).  L >= 4. */
/* Eij_orthotropic(nlm_1, nlm_2, nlm_3, e1, e2, e3, Eij_grain, alpha, n_grain)
 * (src/enhancementfactors.f90:134-188; specfabpy src/specfabpy.f90:436-446 and
 * the *_arr loop :488-500): Sachs homogenization only (alpha unused, as the
 * reference); Eij_grain = (Ebb, Enn, Evv, Env, Ebv, Enb); n_grain = 1 (the
 * reference returns 0/0 otherwise; here an error).  frame: 1 = frames_dev
 * (P, 3, 3) rows e1, e2, e3, 2 = Cartesian. */
int sfg_Eij_orthotropic(sfg_ctx* ctx, int L, const double* nlm1, const double* nlm2, const double* nlm3,
                        const double e1[3], const double e2[3], const double e3[3], const double Eij_grain[6],
                        double alpha, int n_grain, double Eij[6]);
int sfg_Eij_orthotropic_batch(sfg_ctx* ctx, int L, int64_t P, const double* x1_dev, const double* x2_dev,
                              const double* x3_dev, int frame, const double* frames_dev, const double Eij_grain[6],
                              double alpha, int n_grain, double* Eij_dev, void* stream);
/* Evw_orthotropic(nlm_r1, nlm_r2, nlm_r3, v, w, tau, Eij_grain, alpha,
 * n_grain) (src/enhancementfactors.f90:134-160; specfabpy :425-434): one
 * stress state (tau symmetric); batched over the states states_dev (S, 15)
 * as sfg_Evw_tranisotropic_batch, Evw_dev (P, S). */
int sfg_Evw_orthotropic(sfg_ctx* ctx, int L, const double* nlm1, const double* nlm2, const double* nlm3,
                        const double v[3], const double w[3], const double tau[9], const double Eij_grain[6],
                        double alpha, int n_grain, double* Evw);
int sfg_Evw_orthotropic_batch(sfg_ctx* ctx, int L, int64_t P, const double* x1_dev, const double* x2_dev,
                              const double* x3_dev, int nstates, const double* states_dev, const double Eij_grain[6],
                              double alpha, int n_grain, double* Evw_dev, void* stream);
/* Vi_elastic_orthotropic(nlm_1, nlm_2, nlm_3, alpha, Lame_grain, rho, theta,
 * phi) (src/wavepropagation.f90:41-68,98-115; specfabpy :975-984): Voigt
 * average only (alpha unused, as the reference); Lame_grain = (lam11, lam22,
 * lam33, lam23, lam13, lam12, mu1, mu2, mu3).  Vi (ndir, 3) = (qS1, qS2, qP)
 * row-major; batched Vi_dev (P, ndir, 3). */
int sfg_Vi_elastic_orthotropic(sfg_ctx* ctx, int L, const double* nlm1, const double* nlm2, const double* nlm3,
                               double alpha, const double Lame[9], double rho, int ndir, const double* theta,
                               const double* phi, double* Vi);
int sfg_Vi_elastic_orthotropic_batch(sfg_ctx* ctx, int L, int64_t P, const double* x1_dev, const double* x2_dev,
                                     const double* x3_dev, double alpha, const double Lame[9], double rho, int ndir,
                                     const double* theta, const double* phi, double* Vi_dev, void* stream);

/* Discrete grains: the axes b_i, n_i, v_i of N grains, (N, 3) each
 * (ai_orthotropic_discrete, src/moments.f90:386-417, averages over the grains).
 * Eij_orthotropic_discrete(bi, ni, vi, e1, e2, e3, Eij_grain, alpha, n_grain)
 * (src/enhancementfactors.f90:191-266; specfabpy's _arr form :502-514 loops
 * over ensembles): Sachs, normalised by the continuum isotropic fabric.
 * Vi_elastic_orthotropic__discrete(mi, alpha, Lame_grain, rho, theta, phi)
 * (src/wavepropagation.f90:70-96; specfabpy :986-995, mi (3, 3, N) =
 * (axis, xyz, grain)): Voigt, (ndir, 3) = (qS1, qS2, qP) row-major.
 * Batched: E ensembles of N grains, each axis array (E, N, 3) on the device;
 * Eij_dev (E, 6), frames_dev (E, 3, 3) for frame 1; Vi_dev (E, ndir, 3);
 * at most 65535 directions per call. */
int sfg_Eij_orthotropic_discrete(sfg_ctx* ctx, int64_t N, const double* bi, const double* ni, const double* vi,
                                 const double e1[3], const double e2[3], const double e3[3],
                                 const double Eij_grain[6], double alpha, int n_grain, double Eij[6]);
int sfg_Eij_orthotropic_discrete_batch(sfg_ctx* ctx, int64_t E, int64_t N, const double* b_dev, const double* n_dev,
                                       const double* v_dev, int frame, const double* frames_dev,
                                       const double Eij_grain[6], double alpha, int n_grain, double* Eij_dev,
                                       void* stream);
int sfg_Vi_elastic_orthotropic_discrete(sfg_ctx* ctx, int64_t N, const double* bi, const double* ni,
                                        const double* vi, double alpha, const double Lame[9], double rho, int ndir,
                                        const double* theta, const double* phi, double* Vi);
int sfg_Vi_elastic_orthotropic_discrete_batch(sfg_ctx* ctx, int64_t E, int64_t N, const double* b_dev,
                                              const double* n_dev, const double* v_dev, double alpha,
                                              const double Lame[9], double rho, int ndir, const double* theta,
                                              const double* phi, double* Vi_dev, void* stream);
/* Cij_to_Lame_orthotropic (src/elasticities.f90:122-139): (C11, C22, C33, C44,
 * C55, C66, C23, C13, C12) -> (lam11, lam22, lam33, lam23, lam13, lam12, mu1,
 * mu2, mu3).  Host only. */
int sfg_Cij_to_Lame_orthotropic(const double Cij[9], double Lame[9]);

/* Cij_to_Lame_tranisotropic (src/elasticities.f90:107-120): (C11, C33, C55,
 * C12, C13) -> (lam, mu, Elam, Emu, Egam).  Host only. */
int sfg_Cij_to_Lame_tranisotropic(const double Cij[5], double Lame[5]);

/* Vi_elastic_tranisotropic (src/wavepropagation.f90:121-144; specfabpy
 * src/specfabpy.f90:999-1008): phase velocities (qS1, qS2, qP) along the
 * directions (theta[d], phi[d]) (host arrays), grains with Lame parameters
 * Lame[5], Reuss-Voigt weight alpha, density rho.  Vi is (ndir, 3) row-major
 * (the transpose of the reference's Vi(3, ndir)).  Batched: Vi_dev is
 * (P, ndir, 3); a singular Reuss compliance sets status bit
 * SFG_ST_REUSS_SING and NaN velocities (the reference stops). */
int sfg_Vi_elastic_tranisotropic(sfg_ctx* ctx, int L, const double* nlm, double alpha, const double Lame[5],
                                 double rho, int ndir, const double* theta, const double* phi, double* Vi);
int sfg_Vi_elastic_tranisotropic_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, double alpha,
                                       const double Lame[5], double rho, int ndir, const double* theta,
                                       const double* phi, double* Vi_dev, int32_t* status_dev, void* stream);

/* Qnorm_tranisotropic(nlm, alpha, Lame_grain) (src/wavepropagation.f90:146-211,
 * specfabpy :1010-1018): the (1-alpha) Reuss + alpha Voigt acoustic tensor
 * for k = z, row-major 3x3 (P, 9) batched; a singular Reuss compliance
 * gives NaN and SFG_ST_REUSS_SING in status_dev. */
int sfg_Qnorm_tranisotropic(sfg_ctx* ctx, int L, const double* nlm, double alpha, const double Lame[5],
                            double Qnorm[9]);
int sfg_Qnorm_tranisotropic_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, double alpha,
                                  const double Lame[5], double* Qn_dev, int32_t* status_dev, void* stream);

/* Vi_electromagnetic_tranisotropic (src/wavepropagation.f90:217-261): EM
 * phase velocities (S1 fast, S2 slow) for grain permittivities eps_m
 * (parallel to c), eps_t (perpendicular) and permeability mu.  Vi is
 * (ndir, 2); batched Vi_dev is (P, ndir, 2). */
int sfg_Vi_electromagnetic_tranisotropic(sfg_ctx* ctx, int L, const double* nlm, double eps_m, double eps_t,
                                         double mu, int ndir, const double* theta, const double* phi, double* Vi);
int sfg_Vi_electromagnetic_tranisotropic_batch(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, double eps_m,
                                               double eps_t, double mu, int ndir, const double* theta,
                                               const double* phi, double* Vi_dev, void* stream);


/* ---- uneven batches (no counterpart in the reference) ----
 *
 * A plan steps several buckets -- each one truncation L with its own state,
 * kinematics and per-parcel step counts -- through a whole time loop with one
 * graph launch.  Each bucket's desc.steps_dev holds the parcels' remaining
 * step counts (int32, decremented as steps are taken; desc.nsteps is
 * ignored); the time loop is `nchunks` chunks of `chunk` steps, captured once
 * as a HIP graph in which a bucket's chunk c depends only on its chunk c - 1,
 * so buckets run concurrently.  Per-parcel operator weights are computed at
 * creation.  The stepper variant per (L, precision) is the context's
 * autotuned one (sfg_autotune).  Sorting parcels by step count within a
 * bucket keeps wavefronts busy (specfab_amd.batch.UnevenBatch does). */
typedef struct sfg_plan sfg_plan;
int sfg_plan_create(sfg_ctx* ctx, int nbuckets, const sfg_evolve_desc* buckets, int chunk, int nchunks,
                    void* stream, sfg_plan** plan);
int sfg_plan_run(sfg_plan* plan, void* stream);
int sfg_plan_destroy(sfg_plan* plan);

/* ---- learned homogenization (no counterpart in the reference) ----
 *
 * A multilayer perceptron 36 -> 256 -> 256 -> 256 -> 9 (GELU), bf16 weights
 * and activations with FP32 accumulation on the matrix cores, replaces the
 * per-parcel Taylor homogenization (the 6x6 solve of
 * src/homogenizations.f90:145-189) and the elastic-velocity solve:
 *   inputs  Re, Im of n_l^m / n_0^0 for l <= 4 in specfab order (30) and the
 *           unit strain rate D/|D| in Mandel order (6);
 *   outputs log of the Taylor-homogenized enhancement factors T_ij in the
 *           Cartesian frame (E11, E22, E33, E23, E13, E12; Eij_tranisotropic
 *           with alpha = 1) and log of the elastic phase velocities along z,
 *           sorted ascending (slower S, faster S, P).
 * sfg_surrogate_homogenize returns Eij = (1 - alpha) S_ij + alpha T_ij, the
 * linear Sachs/Taylor mix of src/enhancementfactors.f90:23-69, with the Sachs
 * term S_ij in closed form (a linear function of a2 and twelve a4 entries for
 * the axis stress states: no solve).  A parcel whose bound
 * alpha T_ij dT_ij / E_ij on the relative error of any Eij exceeds the
 * tolerance hp->eij_tol (dT_ij: the network's calibrated relative error of
 * T_ij, sfg_mlp_set_error_model) takes the plain homogenization instead
 * (status bit SFG_ST_ERRBOUND).
 * Inputs and outputs are normalised inside the network ((v - mean)/std, set
 * by sfg_mlp_set_normalization).  Parameters are a flat FP32 array of
 * sfg_mlp_nparams() = 143,369 values: W1 (36 x 256, [in][out]), b1, W2
 * (256 x 256), b2, W3, b3, W4 (256 x 9), b4.  Training runs on the GPU (MSE on
 * normalised outputs, Adam on FP32 master weights). */
typedef struct sfg_mlp sfg_mlp;

typedef struct {
    double Eij_grain[2]; /* (Ecc, Eca) */
    double alpha;        /* Sachs-Taylor weight of Eij */
    int n_grain;         /* 1 or -3 */
    double Lame[5];      /* grain elastic parameters (Cij_to_Lame_tranisotropic) */
    double rho;          /* density */
    double vi_alpha;     /* Reuss-Voigt weight of the velocities */
    double eij_tol;      /* surrogate: relative error bound of Eij (<= 0: 1e-3) */
} sfg_homog_params;

typedef struct {
    int64_t n_fallback;  /* parcels given the plain homogenization (envelope or error bound) */
    int64_t n_audit;     /* audited parcels (surrogate kept, plain computed too) */
    double audit_max_err_Eij, audit_mean_err_Eij;  /* max_ij |E - E_plain| / |E_plain| */
    double audit_max_err_Vi, audit_mean_err_Vi;
    int64_t n_bound;     /* of n_fallback: inside the envelope, sent by the error bound */
} sfg_surrogate_stats;

int64_t sfg_mlp_nparams(void);
int sfg_mlp_create(sfg_ctx* ctx, uint64_t seed, sfg_mlp** net);
int sfg_mlp_destroy(sfg_mlp* net);
int sfg_mlp_get_params(sfg_mlp* net, float* params);
int sfg_mlp_set_params(sfg_mlp* net, const float* params);
int sfg_mlp_set_normalization(sfg_mlp* net, const float in_mean[36], const float in_std[36], const float out_mean[9],
                              const float out_std[9]);
int sfg_mlp_get_normalization(sfg_mlp* net, float norm[90]);
/* training envelope: sorted a2 eigenvalues lo[i] <= lambda_i <= hi[i] */
int sfg_mlp_set_envelope(sfg_mlp* net, const double lo[3], const double hi[3]);
/* error model: calibrated bound of the relative error of the network's six
 * Taylor-term outputs (>= 0; all 0, the default, disables the bound) */
int sfg_mlp_set_error_model(sfg_mlp* net, const double taylor_rel_err[6]);
/* one Adam step on B rows: X_dev (B x 36) raw features, T_dev (B x 9) raw
 * targets; *loss (if non-NULL, synchronises) = batch MSE before the update */
int sfg_mlp_train_step(sfg_mlp* net, int64_t B, const float* X_dev, const float* T_dev, float lr, float* loss,
                       void* stream);
/* gradients of the batch loss w.r.t. all parameters (host, nparams), no update */
int sfg_mlp_gradients(sfg_mlp* net, int64_t B, const float* X_dev, const float* T_dev, float* grad, float* loss);
/* inference: X_dev (B x 36) -> Y_dev (B x 9), raw units */
int sfg_mlp_forward(sfg_mlp* net, int64_t B, const float* X_dev, float* Y_dev, void* stream);
/* features (P x 36) of a batch state; L >= 4 */
int sfg_surrogate_features(sfg_ctx* ctx, int L, int64_t P, const double* x_dev, const double* ugrad_dev,
                           float* feat_dev, void* stream);
/* training targets (P x 9) from plain outputs: Eij (P x 6, Cartesian) and
 * Vi (P x 3, along z) */
int sfg_surrogate_targets(int64_t P, const double* Eij_dev, const double* Vi_dev, float* T_dev, void* stream);
/* Eij (P x 6, Cartesian) and sorted Vi (P x 3, along z) of a batch by the
 * network; parcels whose a2 eigenvalues leave the envelope get the plain
 * homogenization (status bit SFG_ST_ENVELOPE), a seeded fraction audit_frac of
 * the others is also computed by the plain method (status bit SFG_ST_AUDIT)
 * and the relative errors are reported in *stats.  The work list stays on the
 * device: with stats == NULL the call is fully stream-ordered on `stream`
 * (no synchronisation); otherwise it synchronises once, at the end. */
int sfg_surrogate_homogenize(sfg_ctx* ctx, sfg_mlp* net, int L, int64_t P, const double* x_dev,
                             const double* ugrad_dev, const sfg_homog_params* hp, double audit_frac, uint64_t seed,
                             double* Eij_dev, double* Vi_dev, int32_t* status_dev, sfg_surrogate_stats* stats,
                             void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SPECFAB_GPU_H */
