/*
 * lbm_hip.h -- C ABI of the MI355X (gfx950) time-stepping core of the
 * PyLBM-Isca linear baroclinic model.
 *
 * The reference's only plugin point is its array-module protocol
 * (pylbm_isca/backend.py:53-77, every hot object keeps ``self.xp``); its hot
 * path is Spectral (pylbm_isca/spectral.py:43-289), Dynamics.tendencies
 * (pylbm_isca/dynamics.py:156-271), Integrator.step
 * (pylbm_isca/integrator.py:264-315), State (pylbm_isca/state.py:26-148) and
 * the LBM.run loop (pylbm_isca/model.py:168-290), batched over ensemble
 * members by pylbm_isca/ensemble.py:41-165.  A device="hip" backend binds the
 * entry points below (ctypes).
 *
 * Conventions
 * - Every entry point returns 0 on success or a negative lbm_status code;
 *   lbm_last_error() gives the message.  No C++ exception crosses the ABI.
 * - Array arguments named d_* are DEVICE pointers (e.g. torch tensors'
 *   data_ptr()) to float64 data in the reference's layouts:
 *     grid 3-D      (B, nlat, nlon, nlev)  C order, latitudes north->south
 *     grid 2-D      (B, nlat, nlon)
 *     spectral 3-D  (B, nmax+1, nmax+1, nlev) complex128, [m, n], zero n<m
 *     spectral 2-D  (B, nmax+1, nmax+1)      complex128
 *   B is a leading member (or batch) axis.
 * - h_* arguments are HOST pointers.
 * - stream is a hipStream_t passed as void* (NULL = the default stream).
 * - The plan owns every table and work buffer; calls on one plan must not
 *   overlap (one plan per stream/thread).
 */
#ifndef LBM_HIP_H
#define LBM_HIP_H

#include <stdint.h>

/* Version of the entry points' signatures below; a binding checks it
 * against lbm_abi_version() of the library it loads. */
#define LBM_ABI_VERSION 5

/* Levels a plan accepts (1 .. LBM_MAX_LEVELS; 20 and 40 run kernels
 * specialised for them, other counts the next register capacity of 16,
 * 32, 48, 64 levels with the count at run time). */
#define LBM_MAX_LEVELS 64

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  LBM_OK = 0,
  LBM_ERR_VALUE = -1,     /* bad argument / shape: Python ValueError           */
  LBM_ERR_RUNTIME = -2,   /* HIP / device failure: Python RuntimeError          */
  LBM_ERR_NONFINITE = -3, /* model blew up: Python FloatingPointError           */
  LBM_ERR_STATE = -4      /* call order (e.g. run before basic state): RuntimeError */
} lbm_status;

/* FP64 or FP32 arithmetic throughout; BF16X3: FP32 with the Legendre
 * contractions as split-BF16x3 on the matrix cores (each FP32 operand split
 * into a BF16 high and low part, three BF16 products in FP32 accumulation).
 * Setup (Gaussian nodes, Legendre recurrences, semi-implicit inverses) is
 * FP64 in every mode. */
typedef enum { LBM_FP64 = 0, LBM_FP32 = 1, LBM_BF16X3 = 2 } lbm_precision;

/* Model description.  Mirrors Spectral(nlon, nlat, nmax, radius)
 * (spectral.py:49-50), VerticalGrid(sigm) (vertical.py:59) and Config
 * (model.py:34-74); unit conversions (hours/days -> s^-1) are done by the
 * caller as in integrator.py:41-118, 243-253. */
typedef struct {
  int32_t nlon, nlat, nmax, nlev;   /* 128, 64, 42, 20 at T42 L20.  nlon: 128
                                       or 384 (specialised FFTs) or a multiple
                                       of 64 up to 1024 with nlon/2 a product
                                       of 2, 3 and 5 (T21 64, T63 192, T85 256,
                                       T106 320 ...); nlat a multiple of 8;
                                       nmax + 1 <= nlon/2; nlev 1 ..
                                       LBM_MAX_LEVELS                         */
  int32_t nmem;                     /* ensemble members E >= 1               */
  int32_t nbasic;                   /* 1 = one shared basic state, or nmem   */
  int32_t precision;                /* lbm_precision                         */
  int32_t strictly_linear;          /* Dynamics(strictly_linear=...)         */
  int32_t inistp;                   /* start-up half steps (model.py:40)     */
  int32_t skip_low_n_vdiff;         /* 1: vor/div vdiff skips n=0,1          */
  double radius;                    /* m                                      */
  double delt;                      /* s                                      */
  double robert;                    /* Robert-Asselin gamma                   */
  const double *h_sigm;             /* nlev+1 half levels, sigm[0]=1 ... 0    */
  /* damping, s^-1 (integrator.py:41-98): */
  const double *h_kv, *h_kd, *h_ktt, *h_kq; /* each (nmax+1) x nlev          */
  /* vertical diffusion nu per level in s^-1, NULL = off (integrator.py:243) */
  const double *h_vdif_vor, *h_vdif_div, *h_vdif_t, *h_vdif_q;
  /* 1: the tangent-linear model about a zero perturbation -- strictly linear
   * tendencies and the linearised pair ps' = ps_bar pi', so the state is
   * the spectral (vor, div, t, q, pi) pair of time levels and the model is
   * exactly linear; required by lbm_adjoint_run.  Not a reference option
   * (the reference's strictly_linear keeps ps' = ps_bar (exp(pi') - 1),
   * state.py:77-83). */
  int32_t tangent_linear;
} lbm_desc;

typedef struct lbm_plan lbm_plan;

/* Builds the Gaussian grid, FP64 associated-Legendre tables
 * (spectral.py:57-154), the Simmons-Burridge operators (vertical.py:59-129),
 * the vertical-diffusion operator (integrator.py:182-206) and allocates
 * every device buffer. */
int lbm_plan_create(const lbm_desc *desc, lbm_plan **out);
int lbm_plan_destroy(lbm_plan *plan);

/* Host copies of setup arrays: "mu", "gw", "coslat", "lat", "lon", "sig",
 * "dsig", "alpha", "cpg", "w3", "wP", "dsigm", "ww" (nlev*nlev), "dlog3",
 * "sigm", "tbar" (nbasic*nlev, after lbm_set_basic_state), "minv"
 * ((inistp+2)*nbasic*(nmax+1)*nlev*nlev: one block per start-up dt, then
 * delt, then the free dt of lbm_prepare_dt), "vinv" ((inistp+2)*4*nlev*nlev),
 * "mmat" (nbasic*(nmax+1)*nlev*nlev: M(n) of the free dt before inversion,
 * SemiImplicit.m of integrator.py:149-158), "lop" (nlev*nlev).
 * n is the capacity of h_out in doubles; returns the count written. */
int lbm_plan_get(const lbm_plan *plan, const char *name, double *h_out, int64_t n);

/* Spectral transforms on caller buffers (spectral.py:186-276).  B members
 * of nk columns each; complex buffers are interleaved re/im float64.
 *   lbm_grid_to_wave   : d_g (B,J,I,nk)            -> d_w (B,M,M,nk)
 *   lbm_wave_to_grid   : d_w (B,M,M,nk)            -> d_g (B,J,I,nk)
 *   lbm_uv_to_vordiv   : d_u, d_v (B,J,I,nk)       -> d_vor, d_div
 *   lbm_psichi_to_uv   : d_psi, d_chi (B,M,M,nk)   -> d_u, d_v (grid)
 *   lbm_hgrad          : d_w (B,M,M,nk)            -> d_gx, d_gy (grid) */
int lbm_grid_to_wave(lbm_plan *plan, const double *d_g, double *d_w, int B, int nk, void *stream);
int lbm_wave_to_grid(lbm_plan *plan, const double *d_w, double *d_g, int B, int nk, void *stream);
int lbm_uv_to_vordiv(lbm_plan *plan, const double *d_u, const double *d_v, double *d_vor,
                     double *d_div, int B, int nk, void *stream);
int lbm_psichi_to_uv(lbm_plan *plan, const double *d_psi, const double *d_chi, double *d_u,
                     double *d_v, int B, int nk, void *stream);
int lbm_hgrad(lbm_plan *plan, const double *d_w, double *d_gx, double *d_gy, int B, int nk,
              void *stream);

/* Transposes of the five transforms above (no counterpart in the reference,
 * which has no adjoint): for the Euclidean inner products of grid values
 * and of the re/im parts of the n >= m coefficients, <T x, y> = <x, T^T y>.
 * Same layouts, inputs and outputs exchanged:
 *   lbm_grid_to_wave_adj : d_w (B,M,M,nk)            -> d_g (grid)
 *   lbm_wave_to_grid_adj : d_g (B,J,I,nk)            -> d_w (spectral)
 *   lbm_uv_to_vordiv_adj : d_vor, d_div (spectral)   -> d_u, d_v (grid)
 *   lbm_psichi_to_uv_adj : d_u, d_v (grid)           -> d_psi, d_chi
 *   lbm_hgrad_adj        : d_gx, d_gy (grid)         -> d_w (spectral) */
int lbm_grid_to_wave_adj(lbm_plan *plan, const double *d_w, double *d_g, int B, int nk,
                         void *stream);
int lbm_wave_to_grid_adj(lbm_plan *plan, const double *d_g, double *d_w, int B, int nk,
                         void *stream);
int lbm_uv_to_vordiv_adj(lbm_plan *plan, const double *d_vor, const double *d_div, double *d_u,
                         double *d_v, int B, int nk, void *stream);
int lbm_psichi_to_uv_adj(lbm_plan *plan, const double *d_u, const double *d_v, double *d_psi,
                         double *d_chi, int B, int nk, void *stream);
int lbm_hgrad_adj(lbm_plan *plan, const double *d_gx, const double *d_gy, double *d_w, int B,
                  int nk, void *stream);

/* Basic state(s): nbasic states, u,v,t,q (nbasic,J,I,L), ps (nbasic,J,I) in
 * hPa.  Computes the BasicStateTerms of dynamics.py:54-93 on the device,
 * the reference temperature, tmtx/gg (vertical.py:139-193) and the FP64
 * semi-implicit matrices M(n) and their inverses for every start-up dt and
 * delt (integrator.py:130-163), plus the vertical-diffusion inverses
 * (integrator.py:216-228). */
int lbm_set_basic_state(lbm_plan *plan, const double *d_u, const double *d_v,
                        const double *d_t, const double *d_q, const double *d_ps,
                        void *stream);

/* Temperature forcing frc_t in K/s, (nfrc,J,I,L) with nfrc = 1 (shared) or
 * nmem; NULL clears it.  The forcing is time constant (ensemble.py:109-121,
 * forcing.py:71-96).  The buffer is copied. */
int lbm_set_forcing(lbm_plan *plan, const double *d_frc_t, int nfrc, void *stream);

/* Initial perturbation, set_initial_perturbation (model.py:142-148): grid
 * fields (nmem,J,I,L) and ps' (nmem,J,I) in hPa for BOTH time levels; NULL
 * pointers mean zero.  Resets the clock. */
int lbm_set_state(lbm_plan *plan, const double *d_u, const double *d_v, const double *d_t,
                  const double *d_q, const double *d_ps, void *stream);

/* Spectral initial perturbation for BOTH time levels: d_vor, d_div, d_t,
 * d_q (nmem,M,M,L) and d_pi = ln ps perturbation (nmem,M,M), complex128;
 * NULL = zero.  The spectral form of set_initial_perturbation
 * (model.py:142-148) followed by State.to_spectral (state.py:86-92): the
 * state is taken as band-limited, so no raw grid step follows.  Resets the
 * clock and the response sum. */
int lbm_set_spectral_state(lbm_plan *plan, const double *d_vor, const double *d_div,
                           const double *d_t, const double *d_q, const double *d_pi,
                           void *stream);

/* set_initial_perturbation exactly (model.py:142-148): each given field
 * replaces both time levels, NULL fields keep each level's current values
 * (all NULL: nothing changes).  Resets the clock like lbm_set_state. */
int lbm_set_state_keep(lbm_plan *plan, const double *d_u, const double *d_v, const double *d_t,
                       const double *d_q, const double *d_ps, void *stream);

/* The filtered previous time level gb (model.py:128-131) on the grid, in
 * lbm_get_state's layouts: ps' as stored, u, v, t, q synthesised from the
 * level's spectral form (the raw grid of lbm_set_state while both levels
 * still hold it). */
int lbm_get_prev_state(lbm_plan *plan, double *d_u, double *d_v, double *d_t, double *d_q,
                       double *d_ps, void *stream);

/* Dynamics.tendencies (dynamics.py:156-271) of a grid state u, v, t, q
 * (nmem,J,I,L) and ps' (nmem,J,I) hPa, any NULL = zero, with the plan's
 * basic state and forcings (lbm_set_forcing's temperature forcing, the
 * steady vorticity / divergence / ln ps forcings; with an MJO forcing the
 * field of the next dynamics call): the spectral tendencies of vor, div,
 * t, q (nmem,M,M,L) and pi (nmem,M,M) complex128, any output NULL.  The
 * run's state is untouched.  Not in the tangent-linear model. */
int lbm_tendencies(lbm_plan *plan, const double *d_u, const double *d_v, const double *d_t,
                   const double *d_q, const double *d_ps, double *d_tvor, double *d_tdiv,
                   double *d_tt, double *d_tq, double *d_tpi, void *stream);

/* SemiImplicit.prepare + VerticalDiffusion.prepare (integrator.py:130-163,
 * 216-228) for a free time step dt: FP64 M(n) and its inverse per basic
 * state, and the vertical-diffusion inverses, in the plan's free matrix
 * set.  Needs the basic state. */
int lbm_prepare_dt(lbm_plan *plan, double dt);

/* Integrator.step (integrator.py:264-315): from the spectral level at t-dt
 * (d_p*), the explicit tendencies (d_t*) and dt, the level at t+dt (d_n*);
 * vor, div, t, q (nmem,M,M,L), pi (nmem,M,M) complex128.  A dt that is none
 * of the start-up steps, delt or the last prepared one is prepared first.
 * No Robert-Asselin filter (the caller's State does that, state.py:110-128);
 * the run's state is untouched. */
int lbm_step(lbm_plan *plan, double dt, const double *d_pvor, const double *d_pdiv,
             const double *d_pt, const double *d_pq, const double *d_ppi, const double *d_tvor,
             const double *d_tdiv, const double *d_tt, const double *d_tq, const double *d_tpi,
             double *d_nvor, double *d_ndiv, double *d_nt, double *d_nq, double *d_npi,
             void *stream);

/* Element-wise operators of State / Spectral / VerticalGrid on caller
 * buffers (float64):
 *   lbm_robert_asselin : prev <- cur + gamma (prev - 2 cur + new), n values
 *                        (State.robert_asselin, state.py:110-128)
 *   lbm_lnps           : inverse = 0: out = ln(ps_bar + in) - ln ps_bar;
 *                        1: out = ps_bar (exp(in) - 1); ps_bar has nper values
 *                        repeated along the leading axis (state.py:77-83)
 *   lbm_grid_stats     : h_out[0] = max |x| over finite values, h_out[1] = the
 *                        number of non-finite values (State.max_abs /
 *                        isfinite, state.py:131-142); synchronises the stream
 *   lbm_spec_scale     : (B,M,M,nk) complex times -n(n+1)/a^2 (kind 0) or
 *                        -a^2/(n(n+1)), 0 at n = 0 (kind 1) (spectral.py:202-208)
 *   lbm_hydrostatic    : phi = x WW^T along the level axis of (npts, nlev,
 *                        inner) values, inner = 1 real or 2 complex, with
 *                        WW = R (diag alpha + strictly lower dlog3 columns)
 *                        from d_alpha, d_dlog3 (nlev each, device)
 *                        (VerticalGrid.hydrostatic, vertical.py:196-203) */
int lbm_robert_asselin(int64_t n, double *d_prev, const double *d_cur, const double *d_new,
                       double gamma, void *stream);
int lbm_lnps(int64_t n, int64_t nper, const double *d_in, const double *d_psbar, double *d_out,
             int inverse, void *stream);
int lbm_grid_stats(int64_t n, const double *d_x, double *h_out, void *stream);

/* lbm_grid_stats_noalloc: the same h_out as lbm_grid_stats, bit for bit,
 * without any allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (tight serving and monitoring
 * loops, harnesses that forbid hipMalloc).  Its partial results live in a
 * static device array of 64 KiB + 256 bytes per GPU, held until the process
 * ends: at most 4096 block pairs, and the launch is capped at that many
 * blocks (grid-stride), so every n >= 0 is accepted.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (two host threads,
 * two streams) race and must be serialised by the caller.  Synchronises the
 * stream. */
int lbm_grid_stats_noalloc(int64_t n, const double *d_x, double *h_out, void *stream);
int lbm_spec_scale(lbm_plan *plan, const double *d_in, double *d_out, int B, int nk, int kind,
                   void *stream);
int lbm_hydrostatic(const double *d_in, double *d_out, int64_t npts, int nlev, int inner,
                    const double *d_alpha, const double *d_dlog3, void *stream);

/* Per-member number of full steps (excluding start-up), NULL = all equal to
 * nsteps.  Members with fewer steps stop advancing when they reach theirs. */
int lbm_set_lengths(lbm_plan *plan, const int32_t *h_nsteps, void *stream);

/* Runs the LBM.run loop (model.py:211-253) for the longest member:
 * start-up half steps then full steps; use_graph = 1 captures the step
 * sequence as a hipGraph and replays it.  check_every > 0 polls an on-device
 * finiteness flag every check_every steps (0 = once at the end) and
 * returns LBM_ERR_NONFINITE naming the step, as model.py:245-247 does.
 * nsteps <= 0 uses the lengths of lbm_set_lengths. */
int lbm_run(lbm_plan *plan, int nsteps, int use_graph, int check_every, void *stream);

/* The same run in segments, for history output between them: runs (or
 * continues) the run of lbm_run(plan, nsteps, ...) until full step stop is
 * done (stop = 0: to the end), without restarting the start-up steps; the
 * run continues with the next call for the same nsteps until its end or
 * until a state is set.  Replaces the history branch of the LBM.run loop
 * (model.py:255-277), which hands each output interval to the writer
 * between steps. */
int lbm_run_to(lbm_plan *plan, int nsteps, int stop, int use_graph, int check_every,
               void *stream);

/* Top-of-step history (model.py:150-166, 189-277).  interval_steps > 0
 * accumulates, from the next run on, the trapezoid 0.5 (s_a + s_{a-1}) of
 * the top-of-step samples s (psi, chi, u, v, w, t, z, p) over each output
 * interval of interval_steps full steps: their mean when time_average = 1
 * (Config.time_average), the last trapezoid of the interval otherwise;
 * interval_steps = 0 stops it.  keep_last = 1 keeps the last top-of-step
 * sample of each member's run for lbm_get_diagnostics. */
int lbm_set_history(lbm_plan *plan, int interval_steps, int time_average, int keep_last,
                    void *stream);

/* Output interval d (1-based) of the current run, after lbm_run_to has
 * passed full step d * interval_steps: psi, chi, u, v, w, t, z (nmem, J, I,
 * L) and p (nmem, J, I) float64, the writer.add fields of model.py:265-272
 * (z without the writer's surface-pressure term, output.py:62-63).  The
 * interval's sums are cleared.  Any output may be NULL. */
int lbm_get_history(lbm_plan *plan, int interval, double *d_psi, double *d_chi, double *d_u,
                    double *d_v, double *d_w, double *d_t, double *d_z, double *d_p,
                    void *stream);

/* LBM.diagnostics after a run (model.py:150-166): psi, chi and w of the
 * last step's tendencies (its top-of-step state) and z = WW T / g of the
 * current state, (nmem, J, I, L) float64; u, v, t, p are lbm_get_state's.
 * Needs keep_last of lbm_set_history before the run. */
int lbm_get_diagnostics(lbm_plan *plan, double *d_psi, double *d_chi, double *d_w, double *d_z,
                        void *stream);

/* Time-varying MJO temperature forcing (forcing.py:99-161), interpolated on
 * the device at every step: d_gtf (nf, ndays, J, I, L) float64 K/s, nf = 1
 * (shared) or nmem; the call counter c = call0 + step (start-up steps
 * included, 1-based) holds day 1 while c < spinup, then moves from day it to
 * it + 1 over nt calls.  single = 1: the fields hold float32 values (a
 * file's real*4) and the slope (b - a) / nt is rounded to float32, as the
 * reference does with them.  lbm_set_forcing replaces it. */
int lbm_set_mjo_forcing(lbm_plan *plan, const double *d_gtf, int nf, int ndays, int nt,
                        int spinup, int call0, int single, void *stream);

/* Steady vorticity and divergence forcings d_vor, d_div (nf, J, I, L) and
 * ln ps forcing d_ps (nf, J, I), float64 per second, nf = 1 or nmem, any of
 * them NULL (all NULL: none): their analyses are added to the vorticity,
 * divergence and pi tendencies of every step (model.py:202-205,
 * dynamics.py:189-190,246-250).  The temperature forcing stays
 * lbm_set_forcing's / lbm_set_mjo_forcing's. */
int lbm_set_steady_forcing(lbm_plan *plan, const double *d_vor, const double *d_div,
                           const double *d_ps, int nf, void *stream);

/* The same run advanced by count executed steps (start-up steps included),
 * for a forcing whose .step() changes every call: the caller replaces the
 * forcing (lbm_set_forcing, same member count) before each step, as the
 * LBM.run loop calls mjo_forcing.step() (model.py:222-229). */
int lbm_run_steps(lbm_plan *plan, int nsteps, int count, int use_graph, int check_every,
                  void *stream);

/* Autotuner of the time step's launches (no counterpart in the reference,
 * whose CuPy path has no launch parameters to tune; the choices it replaces
 * are the shapes cuFFT/cuBLAS pick for the rFFTs and per-m matmuls of
 * spectral.py:159-197): inverse-FFT rows per block, forward-FFT rows per
 * block, and the occupancy target (waves per SIMD) of the level Legendre
 * synthesis and of the level tendency analyses, three variants each, and
 * the grid phase (inverse FFT, tendencies, forward FFT) as three launches
 * (0), with the forward FFT inside the tendency kernel with 8 or 4 (FP32),
 * 4 or 2 (FP64) members per block (1, 2), or in one kernel with 4 or 8
 * (FP32), 2 or 4 (FP64) members per block (3, 4; where it applies: 128
 * longitudes and nmax + 1 <= 48; the raw first step after lbm_set_state
 * takes 0), timed on the plan's own
 * buffers and shapes (best of three trials of reps launches; the state the
 * tendency kernel touches is saved and restored) after
 * lbm_set_basic_state; the fastest of each is kept and the graphs are
 * rebuilt.  Variants differ only in block shape, register budget and
 * fusion: results are bit-identical.  report (len bytes, may be NULL)
 * receives the timings and picks as text. */
int lbm_autotune(lbm_plan *plan, int reps, char *report, int64_t len);

/* Apply (lbm_set_tuning, n <= 5; choices past n are 0) or read
 * (lbm_get_tuning, n >= 5) the variant index of each tuned launch in the
 * order above (0..2; 0..4 for the grid phase), e.g. to reuse a cached
 * autotuner result; a new plan holds the built-in choices (0; for the
 * grid phase where k_grid applies, 4 from 5 members (FP64) or 17 (FP32,
 * 3 from 9), else 0: with fewer members k_grid leaves the GPU part
 * empty); lbm_get_tuning returns the count. */
int lbm_set_tuning(lbm_plan *plan, const int32_t *h_choice, int n);
int lbm_get_tuning(const lbm_plan *plan, int32_t *h_choice, int n);

/* Checkpoints for the adjoint of the default model (no counterpart in the
 * reference): every > 0 makes the following runs keep the whole model state
 * (both time levels, ps' of both, the pending update, loop control) before
 * every `every`-th executed step, in device memory (about
 * 2.6 MB x nmem per checkpoint at T42L20 FP64); lbm_adjoint_run recomputes
 * one segment at a time from them, storing `every` steps of trajectory.
 * 0 turns it off and frees the buffers. */
int lbm_set_checkpoints(lbm_plan *plan, int every);

/* Ends the run in progress without finishing it (e.g. after a writer
 * failed between segments); the next run restarts the start-up steps. */
int lbm_run_abort(lbm_plan *plan);

/* Full steps done by the run in progress (0 when none). */
int lbm_run_segment_done(const lbm_plan *plan, int32_t *full_steps);

/* Spectral resize (regrid.py:68-77): d_src (B, src_nmax+1, src_nmax+1, nk)
 * -> d_dst (B, dst_nmax+1, dst_nmax+1, nk) complex128, zero-padded or with
 * the high-n tail dropped.  No plan needed. */
int lbm_spec_resize(const double *d_src, double *d_dst, int B, int src_nmax, int dst_nmax, int nk,
                    void *stream);

/* Current time level ga in grid space (state.py:94-99 of the last step). */
int lbm_get_state(lbm_plan *plan, double *d_u, double *d_v, double *d_t, double *d_q,
                  double *d_ps, void *stream);

/* Spectral prognostic state (vor, div, t, q (nmem,M,M,L), pi (nmem,M,M))
 * at the current (level 0) or previous filtered (level 1) time level. */
int lbm_get_spectral(lbm_plan *plan, int level, double *d_vor, double *d_div, double *d_t,
                     double *d_q, double *d_pi, void *stream);

/* Number of steps taken since lbm_set_state (start-up included). */
int lbm_steps_taken(const lbm_plan *plan, int64_t *out);

/* Kernel timing with HIP events on the plan's stream (measurement only).
 * lbm_profile(plan, 1) makes every following eager step record an event
 * pair around each of its kernels and accumulate the elapsed times (one
 * stream synchronisation per step); lbm_profile(plan, 0) stops.  Slots, in
 * launch order: 0 synthesis (level fields), 1 synthesis (surface), 2
 * inverse FFT, 3 grid tendencies, 4 forward FFT, 5 analysis (level), 6
 * analysis (surface), 7 spectral step.  lbm_profile_read copies up to n
 * accumulated milliseconds and launch counts and resets them. */
int lbm_profile(lbm_plan *plan, int enable);
int lbm_profile_read(lbm_plan *plan, double *h_ms, int64_t *h_count, int n);

/* Linear response (no counterpart in the reference, which has no adjoint):
 *   J_e = sum_k h_wstep[k] sum_l h_wlev[l] sum_{j,i} h_wgrid[j,i] psi_l(j,i)
 * over the states after step k = 0..nstep-1 of the next lbm_run (start-up
 * steps count), for each member e while it is still advancing; psi_l is the
 * streamfunction (inverse Laplacian of the vorticity) of level l on the
 * grid.  h_wgrid (nlat,nlon), h_wlev (nlev), h_wstep (nstep) are host
 * arrays; h_wgrid = NULL removes the response.  lbm_run accumulates J in
 * FP64; lbm_get_response copies the nmem values to h_j. */
int lbm_set_response(lbm_plan *plan, const double *h_wgrid, const double *h_wlev,
                     const double *h_wstep, int nstep);
int lbm_get_response(lbm_plan *plan, double *h_j);

/* Adjoint (no counterpart in the reference): one backward sweep over the
 * schedule lbm_run(plan, nsteps, ...) follows (start-up half steps,
 * per-member lengths) gives, for every member e, the gradient of J_e of
 * lbm_set_response with respect to that member's spectral initial state of
 * lbm_set_spectral_state (d_gvor, d_gdiv, d_gt, d_gq (nmem,M,M,L), d_gpi
 * (nmem,M,M); complex128 dJ/dRe + i dJ/dIm) and to its temperature forcing
 * (d_gfrc (nmem,J,I,L), K/s).  Any output may be NULL.  Needs a response.
 * tangent_linear = 1: the exactly linear model, nothing stored.  Otherwise
 * (the default model, whose kinetic energy u'u'/2 + v'v'/2 and exp / log
 * pair ps' <-> pi' are not linear, dynamics.py:243-244, state.py:77-83):
 * the gradient at the last run's trajectory, from the checkpoints that run
 * kept (lbm_set_checkpoints before it; the run starts from
 * lbm_set_spectral_state): each segment is recomputed from its checkpoint
 * and swept backward with the transposed kernels about it.  The forward
 * state is left as the run left it.  The forcing must be constant through
 * the run (no MJO forcing, no lbm_set_forcing between its steps):
 * LBM_ERR_STATE otherwise. */
int lbm_adjoint_run(lbm_plan *plan, int nsteps, double *d_gvor, double *d_gdiv, double *d_gt,
                    double *d_gq, double *d_gpi, double *d_gfrc, int use_graph, void *stream);

/* Copies the last error message into buf (always NUL terminated). */
int lbm_last_error(char *buf, int64_t len);

/* LBM_ABI_VERSION of the built library (no counterpart in the reference:
 * lets a binding refuse a library built from other sources). */
int lbm_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif /* LBM_HIP_H */
