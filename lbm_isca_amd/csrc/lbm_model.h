// Device kernels of the linearised model: grid-point tendencies, the
// spectral semi-implicit leapfrog step, basic-state terms and the layout
// conversions between the reference's arrays and the internal ones.
#pragma once
#include "lbm_device.h"

namespace lbm {

constexpr double R_DRY = 287.04;
constexpr double KAPPA = 2.0 / 7.0;
constexpr double EPSVT = 461.50 / 287.04 - 1.0;
constexpr double GRAV = 9.80;  // gravity of the history's z (constants.py:39)
constexpr double OMEGA = 7.2921150e-5;

// Vertical constants, packed in one array (index helpers below).
struct VOff {
  int L;
  __host__ __device__ int dsig() const { return 0; }
  __host__ __device__ int sigm() const { return L; }          // L+1 entries
  __host__ __device__ int alpha() const { return 2 * L + 1; }
  __host__ __device__ int w3() const { return 3 * L + 1; }
  __host__ __device__ int wP() const { return 4 * L + 1; }
  __host__ __device__ int cpg() const { return 5 * L + 1; }
  __host__ __device__ int dlog3() const { return 6 * L + 1; }
  __host__ __device__ int sig() const { return 7 * L + 1; }           // full levels
  __host__ __device__ int size() const { return 8 * L + 1; }
};

// Rows of the basic-state grid block [eb][j][row][I].
struct BRow {
  int L;
  __host__ __device__ int u() const { return 0; }
  __host__ __device__ int v() const { return L; }
  __host__ __device__ int t() const { return 2 * L; }
  __host__ __device__ int q() const { return 3 * L; }
  __host__ __device__ int tv() const { return 4 * L; }
  __host__ __device__ int absvor() const { return 5 * L; }
  __host__ __device__ int div() const { return 6 * L; }
  __host__ __device__ int x5() const { return 7 * L; }
  __host__ __device__ int sdot() const { return 8 * L; }  // L+1 half levels
  __host__ __device__ int pix() const { return 9 * L + 1; }
  __host__ __device__ int piy() const { return 9 * L + 2; }
  __host__ __device__ int psbar() const { return 9 * L + 3; }
  __host__ __device__ int sbu() const { return 9 * L + 4; }  // sum_l dsig_l u_l
  __host__ __device__ int sbv() const { return 9 * L + 5; }  // sum_l dsig_l v_l
  __host__ __device__ int rows() const { return 9 * L + 6; }
};

// Loop control read by the step kernels (graph-replayable: no per-step
// kernel arguments).
struct Ctl {
  int step;      // 0-based step index within the current run
  int raw;       // 1: the next tendency call reads the raw initial grid
  int bad_step;  // first step whose input state was non-finite (INT_MAX = none)
  int pad;
};

static __global__ void k_tick(Ctl *ctl) {
  ctl->step += 1;
  ctl->raw = 0;
}

// ---------------------------------------------------------------------------
// Grid-point tendencies of dynamics.py:156-262, one column per thread,
// levels bottom-up.  Vertical operators are running sums:
//   S_k = sum_{l>=k} dsig (D+A) = S_0 - P_k,  sdot_k = sigm_k S_0 - S_k,
//   (tmtx D)_k = kappa Tref_k (alpha_k D_k + w3_k sum_{l>k} dsig_l D_l),
//   phi_k = sum_{l<k} R dlog3_l x_l + R alpha_k x_k,
// and second-centred vertical advection carries the flux of half level k.
// Outputs per level: fu' fv' ke xu' xv' gt yu' yv' g (' = times 1/(a cos
// lat), the metric of the vector analysis).
// ---------------------------------------------------------------------------
template <class T>
struct ColConst {
  T S0, D0, pix, piy, bpix, bpiy, rc;
  bool full_ke;
};

template <class T>
struct ColState {
  T P = 0, PD = 0, phacc = 0;
  T fl_up = 0, fl_bu = 0, fl_vp = 0, fl_bv = 0, fl_tp = 0, fl_bt = 0, fl_qp = 0, fl_bq = 0;
  T sd_lo = 0, bsd_lo = 0;
  T chk = 0;
};

template <class T>
__device__ inline T vadv_term(T flux_hi, T flux_lo, T f, T s_hi, T s_lo, T rdsig) {
  return (flux_hi - flux_lo - f * (s_hi - s_lo)) * rdsig;
}

// Basic-state values of one level k: u, v, T, q, Tv, absolute vorticity,
// divergence, x5 and sigma-dot of half level k+1; zero above the top.
template <class T>
struct BLev {
  T u = 0, v = 0, t = 0, q = 0, tv = 0, av = 0, dv = 0, x5 = 0, sd = 0;
};

// One level k of the column.  b: the basic state of level k, bn: that of
// level k+1 (u, v, T, q; zero at the top), both loaded a level ahead by the
// caller.  fk: the forcing of level k (0 without one).  (up..qp) at level
// k, (upn..qpn) at k+1 (ignored at the top).
template <class T>
__device__ inline void column_level(int k, int L, const T *__restrict__ vc, const T *__restrict__ tref,
                                    const BLev<T> &b, const BLev<T> &bn, T fk,
                                    const ColConst<T> &cc, ColState<T> &cs, T up, T vp,
                                    T vor, T div, T tp, T qp, T upn, T vpn, T tpn, T qpn, T *out,
                                    T psbar, T pst, T &omega) {
  const VOff vo{L};
  const T eps = (T)EPSVT, kap = (T)KAPPA, rgas = (T)R_DRY, h = (T)0.5;
  const T ds = vc[vo.dsig() + k], rds = (T)1 / ds;
  const T alpha = vc[vo.alpha() + k], w3 = vc[vo.w3() + k], wP = vc[vo.wP() + k];
  const T rcpg = rgas * vc[vo.cpg() + k], dlog3 = vc[vo.dlog3() + k];
  const T tr = tref[k];
  const T bu = b.u, bv = b.v, bt = b.t, bq = b.q;
  const T btv = b.tv, babsv = b.av, bdiv = b.dv, bx5 = b.x5;
  const bool top = (k + 1 == L);
  const T bun = bn.u, bvn = bn.v, btn = bn.t, bqn = bn.q;
  const T bsd_hi = b.sd;
  if (top) upn = vpn = tpn = qpn = 0;
  const T padv = up * cc.bpix + bu * cc.pix + vp * cc.bpiy + bv * cc.piy;
  cs.chk += up + vp + tp + qp;

  const T Pk1 = cs.P + ds * (div + padv), PDk1 = cs.PD + ds * div;
  const T Sk1 = cc.S0 - Pk1;
  const T sd_hi = top ? (T)0 : Pk1 - ((T)1 - vc[vo.sigm() + k + 1]) * cc.S0;
  // flux at half level k+1: s_{k+1} (f_{k+1} + f_k)/2 (0 at the top)
  const T fh_up = top ? 0 : bsd_hi * h * (upn + up);
  const T fh_bu = top ? 0 : sd_hi * h * (bun + bu);
  const T fh_vp = top ? 0 : bsd_hi * h * (vpn + vp);
  const T fh_bv = top ? 0 : sd_hi * h * (bvn + bv);
  const T fh_tp = top ? 0 : bsd_hi * h * (tpn + tp);
  const T fh_bt = top ? 0 : sd_hi * h * (btn + bt);
  const T fh_qp = top ? 0 : bsd_hi * h * (qpn + qp);
  const T fh_bq = top ? 0 : sd_hi * h * (bqn + bq);
  const T va_up = vadv_term(fh_up, cs.fl_up, up, bsd_hi, cs.bsd_lo, rds);
  const T va_bu = vadv_term(fh_bu, cs.fl_bu, bu, sd_hi, cs.sd_lo, rds);
  const T va_vp = vadv_term(fh_vp, cs.fl_vp, vp, bsd_hi, cs.bsd_lo, rds);
  const T va_bv = vadv_term(fh_bv, cs.fl_bv, bv, sd_hi, cs.sd_lo, rds);
  const T va_tp = vadv_term(fh_tp, cs.fl_tp, tp, bsd_hi, cs.bsd_lo, rds);
  const T va_bt = vadv_term(fh_bt, cs.fl_bt, bt, sd_hi, cs.sd_lo, rds);
  const T va_qp = vadv_term(fh_qp, cs.fl_qp, qp, bsd_hi, cs.bsd_lo, rds);
  const T va_bq = vadv_term(fh_bq, cs.fl_bq, bq, sd_hi, cs.sd_lo, rds);

  // temperature (dynamics.py:183, 197-218)
  const T tv = tp + eps * (tp * bq + bt * qp);
  const T x5p = alpha * div + wP * padv + w3 * Sk1;
  const T adiab = -kap * (tv * bx5 + btv * x5p);
  // omega diagnostic -sig (x5' ps_bar + x5_bar ps') [hPa/s] (dynamics.py:265-267)
  omega = -vc[vo.sig() + k] * (x5p * psbar + bx5 * pst);
  const T tdev = bt - tr, tvdev = btv - tr;
  const T xu = up * tdev + bu * tp, xv = vp * tdev + bv * tp;
  const T hadv = tp * bdiv + tdev * div;
  const T tm = kap * tr * (alpha * div + w3 * (cc.D0 - PDk1));
  T gt = adiab + hadv + va_tp + va_bt + tm;
  gt += fk;
  // momentum (dynamics.py:224-244)
  const T fu = vp * babsv + bv * vor - rcpg * (tv * cc.bpix + tvdev * cc.pix) + va_up + va_bu;
  const T fv = -up * babsv - bu * vor - rcpg * (tv * cc.bpiy + tvdev * cc.piy) + va_vp + va_bv;
  const T xvirt = tv - tp;
  const T phiv = cs.phacc + rgas * alpha * xvirt;
  cs.phacc += rgas * dlog3 * xvirt;
  T ke = phiv + up * bu + vp * bv;
  if (cc.full_ke) ke += h * (up * up + vp * vp);
  // humidity (dynamics.py:256-259)
  const T yu = up * bq + bu * qp, yv = vp * bq + bv * qp;
  const T g = qp * bdiv + bq * div + va_qp + va_bq;

  out[0] = fu * cc.rc;
  out[1] = fv * cc.rc;
  out[2] = ke;
  out[3] = xu * cc.rc;
  out[4] = xv * cc.rc;
  out[5] = gt;
  out[6] = yu * cc.rc;
  out[7] = yv * cc.rc;
  out[8] = g;

  cs.P = Pk1;
  cs.PD = PDk1;
  cs.fl_up = fh_up; cs.fl_bu = fh_bu; cs.fl_vp = fh_vp; cs.fl_bv = fh_bv;
  cs.fl_tp = fh_tp; cs.fl_bt = fh_bt; cs.fl_qp = fh_qp; cs.fl_bq = fh_bq;
  cs.sd_lo = sd_hi;
  cs.bsd_lo = bsd_hi;
}

// Surface part of a column: deferred ps' update of the previous step
// (state.py:94-128 on ps': ps'_new = ps_bar (exp(pi) - 1); gb <- ga at
// start-up or the Robert-Asselin filter; ga <- new), continuity
// gt_pi = -S_0 + sum dsig D (dynamics.py:188) and pi of the previous level
// ln(ps_bar + ps'_b) - ln ps_bar = log1p(ps'_b/ps_bar) (state.py:77-79).
template <class T>
__device__ inline void column_surface(T *psa_p, T *psb_p, int mode, T pig, T psbar, T robert,
                                      T S0, T D0, T &gt_pi, T &pi_prev, T &chk, T &ps_t) {
  T psa = *psa_p, psb = *psb_p;
  if (mode != 0) {
    const T psn = psbar * expm1(pig);
    psb = (mode == 1) ? psa : psa + robert * (psb - (T)2 * psa + psn);
    psa = psn;
    *psa_p = psa;
    *psb_p = psb;
  }
  ps_t = psa;
  chk += psa;
  gt_pi = D0 - S0;
  pi_prev = log1p(psb / psbar);
}

// Top-of-step history (model.py:150-166, 189-277).  The reference samples
// psi, chi, u, v, w, t, z, p at the top of every step and averages the
// trapezoid 0.5 (sample_a + sample_{a-1}) over each output interval.  All
// of them but w and p are linear in the spectral vorticity, divergence and
// temperature, so those three are summed in spectral space and synthesised
// once per interval; w (omega) and p (ps') are summed on the grid.  Two
// slots alternate by interval parity: the sample that closes interval d
// also opens d+1.  The last top sample of each member's run is kept for
// LBM.diagnostics (psi, chi and w of the last step's tendencies).
template <class T>
struct HistArgs {
  int mode;           // 0 off, 1 interval means, 2 instantaneous (time_average = false)
  int K;              // full steps per output interval
  int inistp;         // start-up steps: top sample a = step - inistp + 1
  const int *nsteps;  // [E] full steps of each member's run
  cx<T> *spec[2];     // [s][3][L][E] vor, div, t sums, or null
  T *grid[2];         // [e][j][L+1][I] omega levels, then ps', or null
  cx<T> *last_spec;   // [s][2][L][E] vor, div of the last top sample, or null
  T *last_grid;       // [e][j][L][I] omega of the last top sample, or null
};

// (slot, weight) pairs of top sample a (0: the last start-up step; a >= 1:
// full step a) of a member with nst full steps; returns their number.
// mode 1: interval d = ceil(a/K) gets the trapezoid
// 0.5 s_{(d-1)K} + s_{(d-1)K+1} + ... + 0.5 s_{dK}; mode 2: the
// instantaneous 0.5 (s_{dK} + s_{dK-1}).
template <class T>
__device__ inline int hist_weights(int mode, int K, int a, int nst, int *slot, T *w) {
  if (mode == 0 || a < 0 || a > nst) return 0;
  if (mode == 1) {
    if (a == 0) {
      slot[0] = 1;
      w[0] = (T)0.5;
      return 1;
    }
    const int d = (a + K - 1) / K;
    slot[0] = d & 1;
    if (a % K != 0) {
      w[0] = (T)1;
      return 1;
    }
    w[0] = (T)0.5;
    slot[1] = (d + 1) & 1;
    w[1] = (T)0.5;
    return 2;
  }
  int n = 0;
  if ((a + 1) % K == 0) {
    slot[n] = ((a + 1) / K) & 1;
    w[n++] = (T)0.5;
  }
  if (a >= 1 && a % K == 0) {
    slot[n] = (a / K) & 1;
    w[n++] = (T)0.5;
  }
  return n;
}

template <class T>
struct TendArgs {
  const T *gin;     // [e][j][6L+6][I] synthesised state on the grid
  const T *raw;     // [e][j][4L][I] raw initial (u,v,t,q) or null
  const T *bsg;     // [eb][j][9L+6][I]
  const T *frc;     // [ef][j][L][I] or null
  T *gout;          // [e][j][9L+2][I]
  T *psa, *psb;     // [e][j][I] ps' at t and at t-dt (filtered)
  const int *pend;  // [e] deferred update of ps': 0 none, 1 assign, 2 RA filter
  const T *vc;      // vertical constants (VOff)
  const T *tref;    // [eb][L]
  const T *rcos;    // [J] 1/(a cos lat)
  Ctl *ctl;
  int E, J, I, L, nbasic, nfrc, strictly_linear;
  int linear;       // tangent-linear mode: no ps' grid state (pi' is spectral)
  T robert;
  ActiveMask am;
  HistArgs<T> h;
  // trajectory of this step for the adjoint of the default model, or null:
  // [e][j][2L+2][I] = u, v of every level, ps' of both time levels after
  // the deferred update
  T *traj;
  // k_tend<T, true>: the Fourier columns of the 9L + 2 tendency rows
  // (ColMap {E, L, 9, 2}, C columns, M coefficients) instead of gout
  cx<T> *fout;
  int C, M;
  int ne;  // members launched (k_tend<T, true>: tiles of 8 cover them)
  // k_grid: the synthesised Fourier columns (ColMap {E, L, 6, 6}) instead
  // of gin
  const cx<T> *fin;
};

// Members per block of k_tend<T, true>, tuning variant v = 1 or 2: 8 or 4
// (FP32), 4 or 2 (FP64, whose tendency registers do not fit the 128 per
// lane of 1024-thread blocks).
template <class T>
constexpr int tend_fft_members(int v) { return (sizeof(T) == 4 ? 8 : 4) / v; }

// The forward FFT of nf fields x MB members of tendency rows of one
// latitude, held by the MB x 128 threads of a k_tend<T, true> block
// (thread = (member m, longitude i)), through the LDS rows sh[f * MB + m]:
// 8 threads per row (k_fft's stages for 128 points), so each store of a
// coefficient by a wave covers MB consecutive Fourier columns of a field;
// rows of inactive members are skipped.  Every thread of the block calls it.
template <class T, int MB>
__device__ inline void fft_tend_rows(cx<T> (*sh)[72], const T (&vals)[9], int nf, int m, int i,
                                     bool act, unsigned live8, cx<T> *__restrict__ fout,
                                     const int (&col0)[9], int J, int j, int C, int M) {
  __syncthreads();  // the previous rows' last stage has read sh
#pragma unroll
  for (int f = 0; f < 9; ++f)
    if (f < nf && act) reinterpret_cast<T *>(sh[f * MB + m])[i] = vals[f];
  __syncthreads();
  const int rl = threadIdx.x >> 3, tt = threadIdx.x & 7;
  const int rf = rl / MB, rm = rl - rf * MB;
  const bool live = rf < nf && ((live8 >> rm) & 1u);
  cx<T> v[8];
  if (live) fft_a<T, 128, true>(sh[rl], v, tt, twiddles<T, 128>());
  __syncthreads();  // the staged row is overwritten by the first pass
  if (live) fft_a_store<T, 128>(v, sh[rl], tt);
  __syncthreads();
  if (live) fft_b<T, 128>(sh[rl], v, tt);
  __syncthreads();  // the scratch takes Z in natural order
  if (live) {
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) sh[rl][tt + 8 * k1] = v[k1];
  }
  __syncthreads();
  int c = 0;
#pragma unroll
  for (int f = 0; f < 9; ++f)
    if (f == rf) c = col0[f];
  if (live)
    fft_c<T, 128>(sh[rl], fout + fidx(0, j, c + rm, J, C), (size_t)J * fpad(C), M, tt, 0,
                  twiddles<T, 128>());
}

// Tendencies on grid rows, one thread per column, one pass over the levels
// with S_0 and sum dsig D from the synthesised column sums.  The first
// step after a raw (not band-limited) initial perturbation reads the raw
// u, v, T, q and forms S_0 in a first pass over the column, as the
// reference does on its grid.  gin rows: U V vor div T Q (L each), then
// d/dx pi, d/dy pi, pi, sum u, sum v, sum D.
// FFT (128 longitudes): one block per (group of MB members, latitude),
// MB x 128 threads; each level's nine tendency rows of the MB members (and
// first the two surface rows) go through the forward FFT in the block to
// the Fourier columns of a.fout, the values and arithmetic of k_fft; the
// grid rows are never stored.  Members past the launch or finished compute
// on clamped addresses and store nothing.
template <class T, bool FFT = false, int MB = 1>
__global__ __launch_bounds__(FFT ? MB * 128 : 128) void k_tend(TendArgs<T> a) {
  __shared__ cx<T> sh[FFT ? 9 * MB : 1][72];
  int e, j, i, m = 0, e0 = 0;
  bool act = true;
  unsigned live8 = 0;  // FFT: bit m set for the block's active members
  if constexpr (FFT) {
    const int tile = blockIdx.x / a.J;
    j = blockIdx.x - tile * a.J;
    m = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 7);  // wave-uniform member
    i = threadIdx.x & 127;
    e0 = tile * MB;  // the block's first member: its Fourier columns start there
    e = e0 + m;
#pragma unroll
    for (int q = 0; q < MB; ++q) {
      const int eq = tile * MB + q;
      if (eq < a.ne && a.am(eq)) live8 |= 1u << q;
    }
    act = (live8 >> m) & 1u;
    if (e >= a.E) e = a.E - 1;
  } else {
    // blocks of blockDim.x (128 or 64) longitudes:
    // blockIdx = (e * J + j) * (I / blockDim.x) + chunk
    const int nch = a.I / (int)blockDim.x;
    const int ej = blockIdx.x / nch;
    i = (blockIdx.x - ej * nch) * (int)blockDim.x + threadIdx.x;
    e = ej / a.J;
    j = ej - e * a.J;
    if (!a.am(e)) return;  // a finished member: its tendencies are not used
  }
  const int L = a.L, I = a.I, J = a.J;
  const int eb = a.nbasic > 1 ? e : 0;
  const VOff vo{L};
  const BRow br{L};
  const int Rin = 6 * L + 6, Rout = 9 * L + 2;
  const bool use_raw = a.raw != nullptr && a.ctl->raw;
  // the rows of this block's (member, latitude) through buffer views: the
  // row bases sit in scalar registers, every lane adds its longitude
  const unsigned rowb = (unsigned)(I * sizeof(T)), voff = (unsigned)(i * sizeof(T));
  const T *const tag = nullptr;
  const BufView gi(a.gin + (size_t)(e * J + j) * Rin * I, (unsigned)Rin * rowb);
  const BufView rw(use_raw ? a.raw + (size_t)(e * J + j) * 4 * L * I : nullptr, use_raw ? 4u * L * rowb : 0u);
  const BufView bs(a.bsg + (size_t)(eb * J + j) * br.rows() * I, (unsigned)br.rows() * rowb);
  const bool has_fr = a.frc != nullptr;
  const BufView fr(has_fr ? a.frc + (size_t)((a.nfrc > 1 ? e : 0) * J + j) * L * I : nullptr,
                   has_fr ? (unsigned)L * rowb : 0u);
  const BufView go(a.gout + (size_t)(e * J + j) * Rout * I, (unsigned)Rout * rowb);
  auto G = [&](int row) { return gi.template ld<2>(voff, (unsigned)row * rowb, tag); };
  auto RW = [&](int row) { return rw.ld(voff, (unsigned)row * rowb, tag); };
  auto UP = [&](int k) { return use_raw ? RW(k) : G(k); };
  auto VP = [&](int k) { return use_raw ? RW(L + k) : G(L + k); };
  auto TP = [&](int k) { return use_raw ? RW(2 * L + k) : G(4 * L + k); };
  auto QP = [&](int k) { return use_raw ? RW(3 * L + k) : G(5 * L + k); };
  auto BS = [&](int row) { return bs.ld(voff, (unsigned)row * rowb, tag); };
  ColConst<T> cc;
  cc.pix = G(6 * L);
  cc.piy = G(6 * L + 1);
  const T pig = G(6 * L + 2);
  cc.bpix = BS(br.pix());
  cc.bpiy = BS(br.piy());
  const T psbar = BS(br.psbar());
  cc.rc = a.rcos[j];
  cc.full_ke = !a.strictly_linear;
  if (use_raw) {
    T S0 = 0, D0 = 0;
    for (int k = 0; k < L; ++k) {
      const T d = G(3 * L + k);
      const T padv = UP(k) * cc.bpix + BS(br.u() + k) * cc.pix + VP(k) * cc.bpiy +
                     BS(br.v() + k) * cc.piy;
      const T ds = a.vc[vo.dsig() + k];
      S0 += ds * (d + padv);
      D0 += ds * d;
    }
    cc.S0 = S0;
    cc.D0 = D0;
  } else {
    const T su = G(6 * L + 3), sv = G(6 * L + 4), sd = G(6 * L + 5);
    cc.D0 = sd;
    cc.S0 = sd + cc.bpix * su + cc.bpiy * sv + cc.pix * BS(br.sbu()) + cc.piy * BS(br.sbv());
  }
  ColState<T> cs;
  cs.bsd_lo = BS(br.sdot());
  const size_t pidx = (size_t)(e * J + j) * I + i;
  T gpi, pib, pst;
  if (a.linear) {
    gpi = cc.D0 - cc.S0;
    pib = 0;
    pst = psbar * pig;
  } else {
    // an inactive member (FFT blocks) leaves its ps' as it is
    column_surface<T>(a.psa + pidx, a.psb + pidx, act ? a.pend[e] : 0, pig, psbar, a.robert, cc.S0,
                      cc.D0, gpi, pib, cs.chk, pst);
  }
  T *trj = nullptr;
  if (a.traj && act) {
    trj = a.traj + (size_t)(e * J + j) * (2 * L + 2) * I + i;
    trj[(size_t)(2 * L) * I] = a.linear ? pst : a.psa[pidx];
    trj[(size_t)(2 * L + 1) * I] = a.linear ? (T)0 : a.psb[pidx];
  }
  // history samples of this top-of-step state (uniform per block)
  int hslot[2];
  T hw[2];
  const int top_a = a.ctl->step - a.h.inistp + 1;
  const int nst = a.h.nsteps ? a.h.nsteps[e] : 0;
  const int hn = act ? hist_weights<T>(a.h.mode, a.h.K, top_a, nst, hslot, hw) : 0;
  const bool hlast = act && a.h.last_grid != nullptr && top_a == nst;
  const size_t hrow = (size_t)(e * J + j) * (L + 1) * I + i;
  for (int q = 0; q < hn; ++q) a.h.grid[hslot[q]][hrow + (size_t)L * I] += hw[q] * pst;
  const int EL = a.E * L;
  if constexpr (FFT) {
    const T sv[9] = {gpi, pib, 0, 0, 0, 0, 0, 0, 0};
    const int s0 = e0 + 9 * EL;  // first column of the block, surface fields 0 and 1
    const int c0[9] = {s0, s0 + a.E, 0, 0, 0, 0, 0, 0, 0};
    fft_tend_rows<T, MB>(sh, sv, 2, m, i, act, live8, a.fout, c0, J, j, a.C, a.M);
  } else {
    go.st(voff, (unsigned)(9 * L) * rowb, gpi);
    go.st(voff, (unsigned)(9 * L + 1) * rowb, pib);
  }
  const T *tref = a.tref + (size_t)eb * L;
  // the state, basic state and forcing of level k+1 are in flight while
  // level k is formed
  auto blev = [&](int k) {
    BLev<T> b;
    b.u = BS(br.u() + k);
    b.v = BS(br.v() + k);
    b.t = BS(br.t() + k);
    b.q = BS(br.q() + k);
    b.tv = BS(br.tv() + k);
    b.av = BS(br.absvor() + k);
    b.dv = BS(br.div() + k);
    b.x5 = BS(br.x5() + k);
    b.sd = BS(br.sdot() + k + 1);
    return b;
  };
  auto FR = [&](int k) { return has_fr ? fr.ld(voff, (unsigned)k * rowb, tag) : (T)0; };
  T up = UP(0), vp = VP(0), tp = TP(0), qp = QP(0);
  T vor = G(2 * L), dv = G(3 * L), fk = FR(0);
  BLev<T> bc = blev(0);
  for (int k = 0; k < L; ++k) {
    const bool top = (k + 1 == L);
    const T upn = top ? (T)0 : UP(k + 1), vpn = top ? (T)0 : VP(k + 1);
    const T tpn = top ? (T)0 : TP(k + 1), qpn = top ? (T)0 : QP(k + 1);
    const T vorn = top ? (T)0 : G(2 * L + k + 1), dvn = top ? (T)0 : G(3 * L + k + 1);
    const T fkn = top ? (T)0 : FR(k + 1);
    const BLev<T> bn = top ? BLev<T>{} : blev(k + 1);
    T out[9], omega;
    if (trj) {
      trj[(size_t)k * I] = up;
      trj[(size_t)(L + k) * I] = vp;
    }
    column_level<T>(k, L, a.vc, tref, bc, bn, fk, cc, cs, up, vp, vor, dv, tp, qp, upn, vpn, tpn,
                    qpn, out, psbar, pst, omega);
    vor = vorn;
    dv = dvn;
    fk = fkn;
    bc = bn;
    if constexpr (FFT) {
      // first columns (f, k, tile) of the level fields
      int c0[9];
#pragma unroll
      for (int f = 0; f < 9; ++f) c0[f] = f * EL + k * a.E + e0;
      fft_tend_rows<T, MB>(sh, out, 9, m, i, act, live8, a.fout, c0, J, j, a.C, a.M);
    } else {
#pragma unroll
      for (int f = 0; f < 9; ++f) go.template st<2>(voff, (unsigned)(f * L + k) * rowb, out[f]);
    }
    for (int q = 0; q < hn; ++q) a.h.grid[hslot[q]][hrow + (size_t)k * I] += hw[q] * omega;
    if (hlast) a.h.last_grid[((size_t)(e * J + j) * L + k) * I + i] = omega;
    up = upn;
    vp = vpn;
    tp = tpn;
    qp = qpn;
  }
  if (act && !isfinite(cs.chk)) atomicMin(&a.ctl->bad_step, a.ctl->step);
}

// ---------------------------------------------------------------------------
// The grid phase in one kernel (128 longitudes): inverse FFT of the
// synthesised Fourier columns, the tendencies of k_tend and the forward FFT
// of the tendency rows, so that neither grid block (6L + 6 rows in, 9L + 2
// out per member and latitude) passes through memory.  One block per (group
// of MB members, latitude), MB x 128 threads in two roles that alternate
// between block barriers:
//   column phase: thread (member m, longitude i) forms one level of its
//     column (column_level) from grid rows in LDS and writes the level's
//     nine tendency rows to LDS;
//   row phase: 8 threads per row (k_ifft / k_fft's stages for 128 points,
//     whose exchanges stay within the row's wave) transform the tendency
//     rows of the level just formed to their Fourier columns in memory and
//     the synthesised columns of the level after next to grid rows in LDS.
// Grid rows of two levels are held (levels k and k+1 feed level k's
// vertical advection), so each level costs two block barriers.  The blocks
// of one 8-column Fourier tile and latitude are 8 apart in blockIdx: they
// run on one XCD together and their partial lines meet in its L2.  The raw
// first step after set_initial_perturbation takes k_tend.
// ---------------------------------------------------------------------------

// Orders the LDS accesses of one wave around a point (the LDS serves a
// wave's accesses in issue order; this keeps the compiler from moving them).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Members per block of k_grid at NI longitudes, tuning variant v = 3 or 4
// (384 longitudes: one, or two for FP32; the LDS holds no more).
template <class T, int NI>
constexpr int grid_members(int v) {
  return NI == 128 ? (sizeof(T) == 4 ? 4 : 2) << (v - 3) : (sizeof(T) == 4 && v == 4 ? 2 : 1);
}
// Threads per row slot of k_grid: the row's FftShape threads (8 R) rounded
// up to a power of two, so that no row spans two waves.
template <int NI>
constexpr int grid_row_threads() { return NI == 128 ? 8 : 32; }
// Threads of a k_grid block of MB members: 16 MB row slots, and at least
// one column (longitude) per thread.
template <int NI>
constexpr int grid_threads(int MB) {
  return MB * (NI > 16 * grid_row_threads<NI>() ? NI : 16 * grid_row_threads<NI>());
}
// Fourier coefficients one thread of an inverse row prefetches: six per
// thread, m = tt + 8 R q.
template <int NI>
constexpr int grid_pf_m() { return 6 * 8 * (NI / 128); }

// Inverse FFT of one Fourier column from coefficients X(k), k < M (read
// elsewhere, or staged in sc), into the LDS row sc as NI real samples in
// natural order: thread tt of the row's 8 R threads (one wave; k_ifft's
// stages); then() runs once every X(k) has been read.
template <class T, int NI, class XF, class F>
__device__ inline void grid_ifft_from(XF X, int M, cx<T> *sc, int tt, const cx<T> *tw, F then) {
  constexpr int R = FftShape<T, NI>::R;
  cx<T> v[8];
  ifft_a<T, NI>(X, M, v, tt, 0, tw);
  wave_lds_sync();
  then();
  fft_a_store<T, NI>(v, sc, tt);
  wave_lds_sync();
  ifft_b<T, NI>(sc, v, tt, tw);
  wave_lds_sync();
  if constexpr (R == 1) {
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) sc[tt + 8 * k1] = v[k1];
  } else {
    // radix-R across the sub-transforms through the exchange buffer
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) sc[(tt >> 3) * 64 + (tt & 7) + 8 * k1] = v[k1];
    wave_lds_sync();
    cx<T> o[8];
    ifft_c<T, NI>(sc, o, tt);
    wave_lds_sync();
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) sc[(tt & 7) + 8 * k1 + 64 * (tt >> 3)] = o[k1];
  }
}

// The same from the Fourier column itself (M coefficients at stride ms
// from Fr), staged in sc first.
template <class T, int NI>
__device__ inline void grid_ifft_row(const cx<T> *__restrict__ Fr, size_t ms, int M, cx<T> *sc,
                                     int tt, const cx<T> *tw) {
  for (int m = tt; m < M; m += FftShape<T, NI>::TPR) sc[m] = Fr[(size_t)m * ms];
  wave_lds_sync();
  grid_ifft_from<T, NI>([&](int k) { return sc[k]; }, M, sc, tt, tw, [] {});
}

// Forward FFT of the LDS row sc (NI real samples in natural order) to the
// Fourier column F (M coefficients at stride ms); sc is used as scratch
// (k_fft's stages).
template <class T, int NI>
__device__ inline void grid_fft_row(cx<T> *sc, cx<T> *__restrict__ F, size_t ms, int M, int tt,
                                    const cx<T> *tw) {
  constexpr int R = FftShape<T, NI>::R;
  cx<T> v[8];
  fft_a<T, NI, true>(sc, v, tt, tw);
  wave_lds_sync();
  fft_a_store<T, NI>(v, sc, tt);
  wave_lds_sync();
  fft_b<T, NI>(sc, v, tt);
  wave_lds_sync();
#pragma unroll
  for (int k1 = 0; k1 < 8; ++k1) sc[(tt >> 3) + R * ((tt & 7) + 8 * k1)] = v[k1];
  wave_lds_sync();
  fft_c<T, NI>(sc, F, ms, M, tt, 0, tw);
}

// Loads straight into LDS (global_load_lds): lane l of the wave writes
// element l of the wave-uniform 64-element array dst, FP32 complex values
// as two 4-byte halves (the 64 real parts, then the 64 imaginary parts).
// The data arrives under vmcnt: readers wait with wait_vm0 first.
template <class T>
__device__ __forceinline__ void lds_dma(const cx<T> *src, cx<T> *dst);
template <>
__device__ __forceinline__ void lds_dma<double>(const cx<double> *src, cx<double> *dst) {
  __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)src,
                                   (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
}
template <>
__device__ __forceinline__ void lds_dma<float>(const cx<float> *src, cx<float> *dst) {
  float *d = reinterpret_cast<float *>(dst);
  __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)&src->x,
                                   (__attribute__((address_space(3))) void *)d, 4, 0, 0);
  __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)&src->y,
                                   (__attribute__((address_space(3))) void *)(d + 64), 4, 0, 0);
}
template <class T>
__device__ __forceinline__ cx<T> lds_dma_get(const cx<T> *st, int l) {
  if constexpr (sizeof(T) == 8) {
    return st[l];
  } else {
    const float *f = reinterpret_cast<const float *>(st);
    return mk(f[l], f[64 + l]);
  }
}
// s_waitcnt vmcnt(0) (expcnt and lgkmcnt left open)
__device__ __forceinline__ void wait_vm0() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// SBS: one basic state for every member (nbasic = 1), whose level rows the
// block stages in LDS during the row phase before the level is formed;
// otherwise each column loads its own.  The inverse rows' Fourier
// coefficients of the next level are requested one row phase ahead:
// straight to LDS (a 6 x 64 stage per wave that transforms inverse rows),
// or for FP64 with SBS into registers (M <= grid_pf_m<NI>()).
template <class T, int NI, int MB, bool SBS>
__global__ __launch_bounds__(grid_threads<NI>(MB)) void k_grid(TendArgs<T> a) {
  using S = FftShape<T, NI>;
  constexpr int NSIB = 8 / MB;  // blocks per 8-column tile and latitude
  constexpr int NT = grid_threads<NI>(MB);
  constexpr int I = NI;
  constexpr int RT = grid_row_threads<NI>(), TPR = S::TPR;
  // FP64 with SBS keeps the next level's coefficients in registers
  // (measured faster there); the other variants stage them in LDS
  constexpr bool REG = SBS && sizeof(T) == 8 && NI == 128;
  constexpr int NSW = REG ? 1 : (6 * MB * RT + 63) / 64;  // waves holding inverse rows
  // basic-state values a thread loads per level: SBS, its share of the
  // block's nine rows; else the nine values of its column
  constexpr int NBS = SBS ? (9 * I + NT - 1) / NT : 9;
  __shared__ cx<T> rin[2][6][MB][S::LDS];  // synthesised rows of two levels (first: the surface rows)
  __shared__ cx<T> rout[9][MB][S::LDS];    // tendency rows of one level
  __shared__ cx<T> twl[NI];                // the FFT twiddles
  __shared__ T bsl[SBS ? 2 : 1][SBS ? 9 : 1][I];  // SBS: basic-state rows of two levels
  __shared__ cx<T> stg[NSW][6][64];               // next level's coefficients, per wave
  const int J = a.J, L = a.L, E = a.E;
  // blockIdx -> unit u = (tile, group of JP latitudes) and within it the
  // latitude jj < JP and sibling s < NSIB: the JP x NSIB blocks of a unit
  // are 8 apart (one XCD).  A 128-byte line of the Fourier columns holds
  // the 8 columns of a tile at 128 / (8 sizeof(cx<T>)) = JP latitudes.
  constexpr int JP = sizeof(T) == 4 ? 2 : 1;
  constexpr int NSUB = JP * NSIB;
  const int b = blockIdx.x;
  const int u = (b / (8 * NSUB)) * 8 + (b & 7);
  const int sub = (b >> 3) % NSUB;
  const int s = sub % NSIB, jj = sub / NSIB;
  const int ntile = (a.ne + 7) >> 3, nju = J / JP;
  if (u >= ntile * nju) return;
  const int tile = u / nju, j = (u - tile * nju) * JP + jj;
  const int e0 = tile * 8 + s * MB;
  unsigned live = 0;  // bit q: member e0 + q is launched and active
#pragma unroll
  for (int q = 0; q < MB; ++q) {
    const int eq = e0 + q;
    if (eq < a.ne && a.am(eq)) live |= 1u << q;
  }
  if (live == 0) return;
  const int t = (int)threadIdx.x;
  const int EL = E * L;
  const int Ci = 6 * EL + 6 * E, Co = a.C;
  const size_t msi = (size_t)J * fpad(Ci), mso = (size_t)J * fpad(Co);
  const int M = a.M;
  // row role: row slot rl = (group rq, member rm), thread tt of its RT (the
  // first TPR work)
  const int rl = t / RT, tt = t % RT;
  const int rq = rl / MB, rm = rl - rq * MB;
  const bool rlive = tt < TPR && ((live >> rm) & 1u);
  const int er = e0 + rm;
  // column role: member m (wave-uniform), longitude i; colt: the thread
  // holds a column
  const bool colt = NT == MB * NI || t < MB * NI;
  const int m = colt ? __builtin_amdgcn_readfirstlane(t / NI) : MB - 1, i = t % NI;
  const bool act = colt && ((live >> m) & 1u);
  const int e = e0 + m < E ? e0 + m : E - 1;
  const int eb = a.nbasic > 1 ? e : 0;
  const BRow br{L};
  const T *bsj = a.bsg + (size_t)(eb * J + j) * br.rows() * I;
  // basic-state row of field f (u v t q tv absvor div x5, then sigma-dot of
  // the half level below) of level lv
  auto bsrow = [&](int f, int lv) { return f < 8 ? f * L + lv : 8 * L + lv + 1; };
  auto bs_load = [&](int lv, T (&st)[NBS]) {
    if (lv >= L) return;
#pragma unroll
    for (int q = 0; q < NBS; ++q) {
      if constexpr (SBS) {
        const int idx = t + q * NT;
        if (idx < 9 * I) st[q] = bsj[(size_t)bsrow(idx / I, lv) * I + idx % I];
      } else {
        st[q] = bsj[(size_t)bsrow(q, lv) * I + i];
      }
    }
  };
  // SBS: the staged values into the level's slot
  auto bs_store = [&](int lv, const T (&st)[NBS]) {
    if constexpr (SBS) {
      if (lv >= L) return;
#pragma unroll
      for (int q = 0; q < NBS; ++q) {
        const int idx = t + q * NT;
        if (idx < 9 * I) bsl[lv & 1][idx / I][idx % I] = st[q];
      }
    }
  };
  const bool has_fr = a.frc != nullptr;
  const T *frj = has_fr ? a.frc + ((size_t)((a.nfrc > 1 ? e : 0) * J + j) * L) * I + i : nullptr;
  auto FR = [&](int k) { return has_fr && k < L ? frj[(size_t)k * I] : (T)0; };
  // REG: the six coefficients m = tt + TPR q of this thread's inverse row
  // (six named registers, not an array: an array carried around the level
  // loop went to scratch; the loads read coefficient min(m, M - 1))
  cx<T> cf0, cf1, cf2, cf3, cf4, cf5;
  // this thread's coefficients of its inverse row (group rq < 6) at level
  // il: REG into cf0..cf5, else into its wave's stage, coefficient
  // m = tt + TPR q < M at [q][lane]
  auto coef_dma = [&](int il) {
    if (REG && rlive && rq < 6 && il < L) {
      const cx<T> *Fr = a.fin + fidx(0, j, rq * EL + il * E + er, J, Ci);
      auto ld = [&](int m) { return Fr[(size_t)(m < M ? m : M - 1) * msi]; };
      cf0 = ld(tt);
      cf1 = ld(tt + TPR);
      cf2 = ld(tt + 2 * TPR);
      cf3 = ld(tt + 3 * TPR);
      cf4 = ld(tt + 4 * TPR);
      cf5 = ld(tt + 5 * TPR);
    }
    if (!REG && rlive && rq < 6 && il < L) {
      const cx<T> *Fr = a.fin + fidx(0, j, rq * EL + il * E + er, J, Ci) + (size_t)tt * msi;
      const int w = __builtin_amdgcn_readfirstlane(t >> 6);
#pragma unroll
      for (int q = 0; q < 6; ++q)
        if (tt + TPR * q < M) lds_dma<T>(Fr + (size_t)(TPR * q) * msi, stg[w][q]);
    }
  };
  // row phase: inverse rows (groups 0-5) of level il into its slot, when
  // il < L; forward rows (groups 6..) of the tendency rows of level fl
  // (fl < 0: the two surface rows)
  auto row_phase = [&](int fl, int il) {
    if (!rlive) return;
    if (rq < 6) {
      if (il < L) {
        // level il's coefficients landed in the stage before the barrier
        // (wait_vm0 at the end of the column phase); the next level's are
        // requested once they are read
        cx<T> *sc = rin[(il + 1) & 1][rq][rm];
        if constexpr (REG) {
          if (tt < M) sc[tt] = cf0;
          if (tt + TPR < M) sc[tt + TPR] = cf1;
          if (tt + 2 * TPR < M) sc[tt + 2 * TPR] = cf2;
          if (tt + 3 * TPR < M) sc[tt + 3 * TPR] = cf3;
          if (tt + 4 * TPR < M) sc[tt + 4 * TPR] = cf4;
          if (tt + 5 * TPR < M) sc[tt + 5 * TPR] = cf5;
          coef_dma(il + 1);
          wave_lds_sync();
          grid_ifft_from<T, NI>([&](int k) { return sc[k]; }, M, sc, tt, twl, [] {});
        } else {
          const int w = __builtin_amdgcn_readfirstlane(t >> 6), rb = t & (64 - RT);
          grid_ifft_from<T, NI>(
              [&](int k) { return lds_dma_get<T>(stg[w][k / TPR], rb + k % TPR); }, M, sc, tt, twl,
              [] {});
          coef_dma(il + 1);
        }
      }
    } else {
      const int f = rq - 6;
      if (f < (fl < 0 ? 2 : 9)) {
        const int col = (fl < 0 ? 9 * EL + f * E : f * EL + fl * E) + er;
        grid_fft_row<T, NI>(rout[f][rm], a.fout + fidx(0, j, col, J, Co), mso, M, tt, twl);
      }
    }
  };

  // basic state of levels 0 and 1 (SBS: staged; else held per column), the
  // forcing of level 0, the twiddles; the surface rows (slot 0) and level 0
  // (slot 1)
  T bc[NBS], bn[NBS];
  bs_load(0, bc);
  bs_load(1, bn);
  T fk = FR(0);
  for (int q = t; q < NI; q += NT) twl[q] = twiddles<T, NI>()[q];
  bs_store(0, bc);
  bs_store(1, bn);
  __syncthreads();
  if (rlive && rq < 12) {
    const int g = rq < 6 ? rq : rq - 6;
    const int col = (rq < 6 ? 6 * EL + g * E : g * EL) + er;
    grid_ifft_row<T, NI>(a.fin + fidx(0, j, col, J, Ci), msi, M, rin[rq < 6 ? 0 : 1][g][rm], tt,
                         twl);
  }
  coef_dma(1);
  __syncthreads();

  const VOff vo{L};
  auto LR = [&](const cx<T> *row) { return reinterpret_cast<const T *>(row)[i]; };
  auto WR = [&](cx<T> *row, T v) { reinterpret_cast<T *>(row)[i] = v; };
  auto BS = [&](int row) { return bsj[(size_t)row * I + i]; };
  ColConst<T> cc;
  ColState<T> cs;
  T psbar = 0, pst = 0;
  T *trj = nullptr;
  int hslot[2] = {0, 0};
  T hw[2] = {0, 0};
  int hn = 0;
  bool hlast = false;
  const size_t hrow = (size_t)(e * J + j) * (L + 1) * I + i;
  const T *tref = a.tref + (size_t)eb * L;
  if (colt) {
    cc.pix = LR(rin[0][0][m]);
    cc.piy = LR(rin[0][1][m]);
    const T pig = LR(rin[0][2][m]);
    cc.bpix = BS(br.pix());
    cc.bpiy = BS(br.piy());
    psbar = BS(br.psbar());
    cc.rc = a.rcos[j];
    cc.full_ke = !a.strictly_linear;
    {
      const T su = LR(rin[0][3][m]), sv = LR(rin[0][4][m]), sd = LR(rin[0][5][m]);
      cc.D0 = sd;
      cc.S0 = sd + cc.bpix * su + cc.bpiy * sv + cc.pix * BS(br.sbu()) + cc.piy * BS(br.sbv());
    }
    cs.bsd_lo = BS(br.sdot());
    const size_t pidx = (size_t)(e * J + j) * I + i;
    T gpi, pib;
    if (a.linear) {
      gpi = cc.D0 - cc.S0;
      pib = 0;
      pst = psbar * pig;
    } else {
      // an inactive member leaves its ps' as it is
      column_surface<T>(a.psa + pidx, a.psb + pidx, act ? a.pend[e] : 0, pig, psbar, a.robert,
                        cc.S0, cc.D0, gpi, pib, cs.chk, pst);
    }
    if (a.traj && act) {
      trj = a.traj + (size_t)(e * J + j) * (2 * L + 2) * I + i;
      trj[(size_t)(2 * L) * I] = a.linear ? pst : a.psa[pidx];
      trj[(size_t)(2 * L + 1) * I] = a.linear ? (T)0 : a.psb[pidx];
    }
    // history samples of this top-of-step state
    const int top_a = a.ctl->step - a.h.inistp + 1;
    const int nst = a.h.nsteps ? a.h.nsteps[e] : 0;
    hn = act ? hist_weights<T>(a.h.mode, a.h.K, top_a, nst, hslot, hw) : 0;
    hlast = act && a.h.last_grid != nullptr && top_a == nst;
    for (int q = 0; q < hn; ++q) a.h.grid[hslot[q]][hrow + (size_t)L * I] += hw[q] * pst;
    WR(rout[0][m], gpi);
    WR(rout[1][m], pib);
  }
  if constexpr (!REG) wait_vm0();  // the stage of level 1
  __syncthreads();
  row_phase(-1, 1);  // the surface tendency rows out, level 1 in
  __syncthreads();

  for (int k = 0; k < L; ++k) {
    if (colt) {
      const int sk = (k + 1) & 1, sn = k & 1;  // slots of levels k and k+1
      const bool top = (k + 1 == L);
      BLev<T> bk, bk1;  // basic state of level k; u, v, T, q of level k+1
      if constexpr (SBS) {
        const int b0 = k & 1, b1 = (k + 1) & 1;
        bk.u = bsl[b0][0][i];
        bk.v = bsl[b0][1][i];
        bk.t = bsl[b0][2][i];
        bk.q = bsl[b0][3][i];
        bk.tv = bsl[b0][4][i];
        bk.av = bsl[b0][5][i];
        bk.dv = bsl[b0][6][i];
        bk.x5 = bsl[b0][7][i];
        bk.sd = bsl[b0][8][i];
        if (!top) {
          bk1.u = bsl[b1][0][i];
          bk1.v = bsl[b1][1][i];
          bk1.t = bsl[b1][2][i];
          bk1.q = bsl[b1][3][i];
        }
      } else {
        bk = BLev<T>{bc[0], bc[1], bc[2], bc[3], bc[4], bc[5], bc[6], bc[7], bc[8]};
        if (!top) {
          bk1.u = bn[0];
          bk1.v = bn[1];
          bk1.t = bn[2];
          bk1.q = bn[3];
        }
      }
      const T up = LR(rin[sk][0][m]), vp = LR(rin[sk][1][m]);
      const T vor = LR(rin[sk][2][m]), dv = LR(rin[sk][3][m]);
      const T tp = LR(rin[sk][4][m]), qp = LR(rin[sk][5][m]);
      T upn = 0, vpn = 0, tpn = 0, qpn = 0;
      if (!top) {
        upn = LR(rin[sn][0][m]);
        vpn = LR(rin[sn][1][m]);
        tpn = LR(rin[sn][4][m]);
        qpn = LR(rin[sn][5][m]);
      }
      T out[9], omega;
      if (trj) {
        trj[(size_t)k * I] = up;
        trj[(size_t)(L + k) * I] = vp;
      }
      column_level<T>(k, L, a.vc, tref, bk, bk1, fk, cc, cs, up, vp, vor, dv, tp, qp, upn, vpn,
                      tpn, qpn, out, psbar, pst, omega);
#pragma unroll
      for (int f = 0; f < 9; ++f) WR(rout[f][m], out[f]);
      for (int q = 0; q < hn; ++q) a.h.grid[hslot[q]][hrow + (size_t)k * I] += hw[q] * omega;
      if (hlast) a.h.last_grid[((size_t)(e * J + j) * L + k) * I + i] = omega;
    }
    if constexpr (!REG) wait_vm0();  // the stage of level k + 2 (requested in the last row phase)
    __syncthreads();
    // level k's tendency rows out, level k + 2 in; meanwhile the basic
    // state of level k + 2 and this column's forcing of level k + 1 arrive
    T st[NBS];
    bs_load(k + 2, st);
    const T fkn = FR(k + 1);
    row_phase(k, k + 2);
    bs_store(k + 2, st);
    if constexpr (!SBS) {
#pragma unroll
      for (int q = 0; q < NBS; ++q) {
        bc[q] = bn[q];
        bn[q] = st[q];
      }
    }
    fk = fkn;
    __syncthreads();
  }
  if (act && !isfinite(cs.chk)) atomicMin(&a.ctl->bad_step, a.ctl->step);
}

// ---------------------------------------------------------------------------
// Spectral step, integrator.py:264-315 fused with the Robert-Asselin filter
// (state.py:110-128, applied to the spectral fields it is linear in) and
// the start-up schedule of model.py:211-219.  One thread per (s, member),
// all levels in registers.
// ---------------------------------------------------------------------------
template <class T>
struct StepArgs {
  cx<T> *sa3, *sa2, *sb3;        // [s][4][L][E], [s][3][E] (pi, sum vor, sum div), [s][4][L][E]
  cx<T> *sb2;                    // [s][3][E] field 0: pi at t-dt (tangent-linear mode) or null
  const cx<T> *tend3, *tend2;    // [s][4][L][E], [s][2][E] (wt_pi, pi_prev)
  const T *kv, *kd, *ktt, *kq;   // [M][L]
  const T *minv;                 // [nset][nbasic][M][L][L]
  const T *vinv;                 // [nset][4][L][L]
  const T *vc, *tref, *lap;      // VOff, [nbasic][L], [M]
  const short *n_of_s;           // [nspec]
  Ctl *ctl;
  int *pend;                     // [E]
  const int *nsteps;             // [E] full steps per member
  int E, M, nspec, nbasic, inistp, vmask, skip_low_n;
  int L;                         // levels
  int ecnt;                      // members covered by the launch (<= E)
  T delt, robert;
  // >= 0: one Integrator.step of the given dt with matrix set fixed_set on
  // caller buffers (integrator.py:264-315): every member advances, no
  // start-up schedule, no Robert-Asselin filter, no deferred ps' update and
  // the global means are kept
  int fixed_set;
  T fixed_dt;
};

// Read-only pointers: through the constant address space when the address
// is wave-uniform, so the compiler may load through the scalar cache (the
// tables are written only between kernels).
template <class T, bool UNIFORM>
struct RoPtr {
  using type = const T *;
};
template <class T>
struct RoPtr<T, true> {
  using type = const T __attribute__((address_space(4))) *;
};

template <class T, int LT, bool EXACT, class P>
__device__ inline void matvec(P A, const cx<T> *x, cx<T> *y, int L) {
#pragma unroll
  for (int r = 0; r < LT; ++r) if (EXACT || r < L) {
    cx<T> acc = zero<T>();
#pragma unroll
    for (int c = 0; c < LT; ++c) if (EXACT || c < L) fma_acc(acc, A[r * L + c], x[c]);
    y[r] = acc;
  }
}

// One thread per (s, member), members fastest across lanes so every field
// access of a wave is contiguous; all levels in registers.  UNI: every wave
// holds one coefficient s and the members share one basic state (launch
// width a multiple of 64, nbasic = 1), so the degree, the damping rows and
// the L x L matrices are wave-uniform and come through the scalar cache
// instead of one vector load per element.
//
// NG = 1 (one basic state, any member count): wave w holds the (m, member)
// pairs of one total wavenumber n, m fastest-but-one and members fastest,
// ceil((n + 1) ecnt / 64) waves per n in increasing n; the per-n matrix,
// damping rows and degree are wave-uniform as under UNI.  (In the order of
// the coefficients, m-major, consecutive lanes would each hold another n
// and every matrix load would gather 64 rows.)  NG = G > 1 (a basic state
// per member, member counts that are multiples of G = 128 bytes / complex
// value: 8 FP64, 16 FP32): the same waves with the pairs ordered (group of
// G members, m, member in the group), so a wave spans G members' matrices
// instead of 64 (each load of one matrix element touches G rows, not 64)
// and its state accesses stay whole 128-byte runs.
//
// LT levels in registers.  EXACT: the plan has LT levels (compile-time
// loops); otherwise it has L = a.L <= LT, the loops run to LT and skip the
// levels past L, whose registers are never read.
// Coefficient s, member e and total wavenumber n of this lane under the NG
// mappings (NG > 0) of k_step and k_step_adj; false for a lane past its
// wave's pairs.
template <int NG>
__device__ __forceinline__ bool ng_lane(int M, int ecnt, int &s, int &e, int &n) {
  // this wave's n: the waves of n' < n come first
  const int w = blockIdx.x;
  int wb = 0;
  n = 0;
  while (n < M - 1) {
    const int nw = ((n + 1) * ecnt + 63) / 64;
    if (w < wb + nw) break;
    wb += nw;
    ++n;
  }
  const int pr = (w - wb) * 64 + (int)threadIdx.x;  // pair (m, member) of n
  if (pr >= (n + 1) * ecnt) return false;
  int mm;
  if constexpr (NG == 1) {
    mm = pr / ecnt;
    e = pr - mm * ecnt;
  } else {
    const int g = pr / NG;  // (group of NG members, m)
    mm = g % (n + 1);
    e = (g / (n + 1)) * NG + pr % NG;
  }
  s = spec_off(mm, M) + n - mm;
  return true;
}

template <class T, int LT, bool UNI = false, bool EXACT = true, int NG = 0>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(LT <= 20 ? 2 : 1)))
void k_step(StepArgs<T> a) {
  const int L = EXACT ? LT : a.L;
  const int E = a.E;
  int s, e, ng_n = 0;
  if constexpr (NG > 0) {
    if (!ng_lane<NG>(a.M, a.ecnt, s, e, ng_n)) return;
  } else {
    const int idx = blockIdx.x * 64 + threadIdx.x;
    s = UNI ? __builtin_amdgcn_readfirstlane(idx / a.ecnt) : idx / a.ecnt;
    e = idx - s * a.ecnt;
  }
  if (s >= a.nspec) return;
  const bool pure = a.fixed_set >= 0;
  const int st = pure ? a.inistp : a.ctl->step;
  const bool active = pure || st < a.inistp + a.nsteps[e];
  const bool oinit = st < a.inistp;
  // blockIdx.y: 0 -> temperature, divergence, pi (the coupled solve);
  // 1 -> vorticity and humidity (independent of it): twice the waves, each
  // with half the dependent chain
  const int role = blockIdx.y;
  if (!pure && s == 0 && role == 0) a.pend[e] = active ? (oinit ? 1 : 2) : 0;
  if (!active) return;
  const T dt = pure ? a.fixed_dt : oinit ? a.delt / (T)(1 << (a.inistp - st)) : a.delt;
  const int set = pure ? a.fixed_set : oinit ? st : a.inistp;
  const T dt2 = (T)2 * dt;
  const int n = NG > 0 ? __builtin_amdgcn_readfirstlane(ng_n)
                       : UNI ? __builtin_amdgcn_readfirstlane(a.n_of_s[s]) : a.n_of_s[s];
  const int eb = (!UNI && NG != 1 && a.nbasic > 1) ? e : 0;
  const VOff vo{L};
  using RU = typename RoPtr<T, true>::type;            // uniform in every launch
  using RP = typename RoPtr<T, UNI || NG == 1>::type;  // uniform: one n, one basic state per wave
  const RU vc = (RU)a.vc;
  const RP tref = (RP)(a.tref + (size_t)eb * L);
  const RP ktt = (RP)(a.ktt + (size_t)n * L), kd = (RP)(a.kd + (size_t)n * L);
  const RP kv = (RP)(a.kv + (size_t)n * L), kq = (RP)(a.kq + (size_t)n * L);
  const size_t EL = (size_t)E * L;
  // Level l of field f of this (s, member) through buffer views based at the
  // wave's first coefficient s_w: a per-lane offset (coefficient within the
  // wave, member) and a wave-uniform level offset, since the level stride E
  // is not a compile-time constant (64-bit addresses per level would not
  // fit in registers).  A wave spans at most 64 coefficients, so the offsets
  // stay below 2^32 bytes; under NG its coefficients (one n) lie up to
  // nspec apart, and the host takes NG only while nspec 4 E L sizeof(cx<T>)
  // < 2^32.  The base is the wave's lowest coefficient: the first lane's
  // (m increasing across the lanes) or, NG > 1 (m restarting with each
  // member group), that of m = 0, s = n.
  const int s_w = NG > 1 ? n : __builtin_amdgcn_readfirstlane(s);
  const size_t wbase = (size_t)s_w * 4 * EL;
  const unsigned voff = (unsigned)(((size_t)(s - s_w) * 4 * EL + e) * sizeof(cx<T>));
  const size_t wbytes = (size_t)(a.nspec - s_w) * 4 * EL * sizeof(cx<T>);
  const unsigned wrange = wbytes < 0xffffffffull ? (unsigned)wbytes : 0xffffffffu;
  const BufView bsa(a.sa3 + wbase, wrange), bsb(a.sb3 + wbase, wrange), btn(a.tend3 + wbase, wrange);
  auto lo = [&](int f, int l) __attribute__((always_inline)) {
    return (unsigned)(((size_t)f * EL + (size_t)l * E) * sizeof(cx<T>));
  };
  const cx<T> *const tag = nullptr;  // selects the precision of BufView::ld
  // tendencies: non-temporal at their last read (streamed once per step)
  const cx<T> wt_pi = a.tend2[(size_t)s * 2 * E + e];
  // pi at t-dt: the analysed log1p(ps'_b/ps_bar), or the spectral level
  // itself in tangent-linear mode
  const cx<T> pi_prev = a.sb2 ? a.sb2[(size_t)s * 3 * E + e] : a.tend2[(size_t)s * 2 * E + E + e];
  const T lapn = ((RP)a.lap)[n];
  const T rg = (T)R_DRY, kap = (T)KAPPA, gam = a.robert;
  cx<T> *s2 = a.sa2 + (size_t)s * 3 * E + e;

  // Level arrays in registers: x, and dbar unless DB_MEM.  Every mat-vec
  // row goes straight to the Robert-Asselin update and to memory.  DB_MEM
  // (FP32, 40 levels, measured faster there only): dbar goes row by row into
  // the divergence-tendency slot (its last read is the rhs) and is read back
  // where used, one level array fewer for a round trip through the cache.
  constexpr bool DB_MEM = sizeof(T) == 4 && LT == 40 && EXACT;
  cx<T> x[LT], dbar[DB_MEM ? 1 : LT];
  const RU vinv = (RU)(a.vinv + (size_t)set * 4 * L * L);
  const bool vd_low = !(a.skip_low_n && n <= 1);
  // The reference keeps winds, not vorticity and divergence, on its grid:
  // their global means (n = 0) never survive a step (a forcing's mean
  // still acts through dbar within the step).  Same here, outside the
  // tangent-linear model.
  const bool drop_mean = n == 0 && a.sb2 == nullptr && !pure;
  // row r of V (or the identity) applied to x, then the filter and the store
  // RD > 0: the filter's two inputs of row r + RD are requested before row
  // r is stored.  The compiler keeps a load after a store it cannot prove
  // disjoint, so with each row's inputs loaded just before its store (RD =
  // 0) every level waits a full memory round trip.  FP32: RD = 3 (measured:
  // 64 members with their own basic states 0.631 -> 0.601 ms per T42L20
  // step); FP64 keeps RD = 0 (rows in flight spill there, and were no
  // faster).
  constexpr int RD = sizeof(T) == 8 ? 0 : 3;
  auto update = [&](int f, RU V, bool apply, cx<T> *sum, bool drop) __attribute__((always_inline)) {
    cx<T> acc = zero<T>();
    cx<T> pa[RD > 0 ? RD : 1], pb[RD > 0 ? RD : 1];  // rows r .. r + RD - 1 of both levels
    if (RD > 0 && !pure) {
#pragma unroll
      for (int q = 0; q < RD; ++q) if (q < LT && (EXACT || q < L)) {
        pa[q] = bsa.ld(voff, lo(f, q), tag);
        pb[q] = bsb.ld(voff, lo(f, q), tag);
      }
    }
#pragma unroll
    for (int r = 0; r < LT; ++r) if (EXACT || r < L) {
      cx<T> y = x[r];
      if (apply) {
        y = zero<T>();
#pragma unroll
        for (int c = 0; c < LT; ++c) if (EXACT || c < L) fma_acc(y, V[r * L + c], x[c]);
      }
      if (drop) y = zero<T>();
      if (RD == 0 && !pure) {
        const cx<T> cur = bsa.ld(voff, lo(f, r), tag);
        bsb.st(voff, lo(f, r), oinit ? cur : cur + gam * (bsb.ld(voff, lo(f, r), tag) - (T)2 * cur + y));
      } else if (!pure) {
        constexpr int R1 = RD > 0 ? RD : 1;
        const cx<T> cur = pa[r % R1], prev = pb[r % R1];
        if (r + RD < LT && (EXACT || r + RD < L)) {
          pa[r % R1] = bsa.ld(voff, lo(f, r + RD), tag);
          pb[r % R1] = bsb.ld(voff, lo(f, r + RD), tag);
        }
        bsb.st(voff, lo(f, r), oinit ? cur : cur + gam * (prev - (T)2 * cur + y));
      }
      bsa.st(voff, lo(f, r), y);
      if (sum) acc = acc + vc[vo.dsig() + r] * y;
    }
    if (sum) *sum = acc;
  };

  if (role == 1) {
    // vorticity
#pragma unroll
    for (int l = 0; l < LT; ++l) if (EXACT || l < L) x[l] = (T)1 / ((T)1 + dt2 * kv[l]) * (bsb.ld(voff, lo(0, l), tag) + dt2 * btn.ld_nt(voff, lo(0, l), tag));
    update(0, vinv, (a.vmask & 1) && vd_low, s2 + E, drop_mean);
    // humidity
#pragma unroll
    for (int l = 0; l < LT; ++l) if (EXACT || l < L)
      x[l] = (T)1 / ((T)1 + dt2 * kq[l]) * (bsb.ld(voff, lo(3, l), tag) + dt2 * btn.ld_nt(voff, lo(3, l), tag));
    update(3, vinv + 3 * L * L, (a.vmask & 8) != 0, nullptr, false);
    return;
  }
  // temperature estimate and its geopotential (integrator.py:287-292)
#pragma unroll
  for (int l = 0; l < LT; ++l) if (EXACT || l < L) x[l] = ((T)1 + dt * ktt[l]) * bsb.ld(voff, lo(2, l), tag) + dt * btn.ld(voff, lo(2, l), tag);
  {
    cx<T> acc = zero<T>();
#pragma unroll
    for (int l = 0; l < LT; ++l) if (EXACT || l < L) {
      const cx<T> wl = x[l];
      x[l] = acc + (rg * vc[vo.alpha() + l]) * wl;
      acc = acc + (rg * vc[vo.dlog3() + l]) * wl;
    }
  }
  const cx<T> wdpix = pi_prev + dt * wt_pi;
#pragma unroll
  for (int l = 0; l < LT; ++l) if (EXACT || l < L) {
    const T ft = (T)1 + dt2 * ktt[l];
    const T gg = rg * tref[l] * vc[vo.cpg() + l];
    const cx<T> phy = x[l] + (ft * gg) * wdpix;
    x[l] = (((T)1 + dt * kd[l]) * ft) * bsb.ld(voff, lo(1, l), tag) + dt * (ft * btn.ld_nt(voff, lo(1, l), tag) - lapn * phy);
  }
  // dbar = M(n)^-1 rhs (integrator.py:165-172)
  cx<T> sdb = zero<T>();
  {
    const RP mi = (RP)(a.minv + (((size_t)set * a.nbasic + eb) * a.M + n) * L * L);
#pragma unroll
    for (int r = 0; r < LT; ++r) if (EXACT || r < L) {
      cx<T> acc = zero<T>();
#pragma unroll
      for (int c = 0; c < LT; ++c) if (EXACT || c < L) fma_acc(acc, mi[r * L + c], x[c]);
      if constexpr (DB_MEM)
        btn.st(voff, lo(1, r), acc);
      else
        dbar[r] = acc;
      sdb = sdb + vc[vo.dsig() + r] * acc;
    }
  }
  auto DB = [&](int l) __attribute__((always_inline)) {
    if constexpr (DB_MEM)
      return btn.ld(voff, lo(1, l), tag);
    else
      return dbar[l];
  };

  // surface pressure (integrator.py:301, 308)
  {
    const cx<T> pin = pi_prev + dt2 * (wt_pi - sdb);
    if (a.sb2) {
      const cx<T> cur = s2[0];
      a.sb2[(size_t)s * 3 * E + e] = oinit ? cur : cur + gam * (pi_prev - (T)2 * cur + pin);
    }
    s2[0] = pin;
  }
  // divergence
#pragma unroll
  for (int l = 0; l < LT; ++l) if (EXACT || l < L) x[l] = (T)2 * DB(l) - bsb.ld(voff, lo(1, l), tag);
  update(1, vinv + 1 * L * L, (a.vmask & 2) && vd_low, s2 + 2 * E, drop_mean);
  // temperature: restore -tmtx dbar (integrator.py:300), damped update
  {
    cx<T> suf = zero<T>();
#pragma unroll
    for (int l = LT - 1; l >= 0; --l) if (EXACT || l < L) {
      const T ft = (T)1 + dt2 * ktt[l];
      const cx<T> db = DB(l);
      const cx<T> tmd = (kap * tref[l]) * (vc[vo.alpha() + l] * db + vc[vo.w3() + l] * suf);
      suf = suf + vc[vo.dsig() + l] * db;
      x[l] = (T)(1) / ft * (bsb.ld(voff, lo(2, l), tag) + dt2 * (btn.ld_nt(voff, lo(2, l), tag) - tmd));
    }
  }
  update(2, vinv + 2 * L * L, (a.vmask & 4) != 0, nullptr, false);
}

// Samples of the top-of-step spectral state (the level the next k_step
// replaces): the interval sums of vorticity, divergence and temperature and
// the last sample (vorticity, divergence) of each member's run.  One thread
// per (s, member), levels in a loop.
template <class T>
__global__ void k_hist_acc(const cx<T> *__restrict__ sa3, HistArgs<T> h, const Ctl *ctl, int E,
                           int L, int nspec, int ecnt) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t s = idx / ecnt, e = idx - s * ecnt;
  if (s >= (size_t)nspec) return;
  const int top_a = ctl->step - h.inistp + 1;
  int slot[2];
  T w[2];
  const int n = hist_weights<T>(h.mode, h.K, top_a, h.nsteps[e], slot, w);
  const bool last = h.last_spec != nullptr && top_a == h.nsteps[e];
  if (n == 0 && !last) return;
  const size_t EL = (size_t)E * L;
  for (int f = 0; f < 3; ++f)
    for (int l = 0; l < L; ++l) {
      const cx<T> x = sa3[(s * 4 + f) * EL + (size_t)l * E + e];
      const size_t dst = (s * 3 + f) * EL + (size_t)l * E + e;
      for (int q = 0; q < n; ++q) h.spec[slot[q]][dst] = h.spec[slot[q]][dst] + w[q] * x;
      if (last && f < 2) h.last_spec[(s * 2 + f) * EL + (size_t)l * E + e] = x;
    }
}

// Column sums sum_l dsig_l x_l of the spectral vorticity and divergence of
// the current level into sa2 fields 1 and 2 (after a state is loaded).
template <class T>
__global__ void k_vsum(const cx<T> *__restrict__ sa3, cx<T> *__restrict__ sa2, const T *__restrict__ vc,
                       int E, int L, int nspec) {
  const size_t tot = (size_t)nspec * E;
  const VOff vo{L};
  const size_t EL = (size_t)E * L;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot;
       q += (size_t)gridDim.x * blockDim.x) {
    const size_t s = q / E, e = q - s * E;
    cx<T> sv = zero<T>(), sd = zero<T>();
    for (int l = 0; l < L; ++l) {
      sv = sv + vc[vo.dsig() + l] * sa3[s * 4 * EL + (size_t)l * E + e];
      sd = sd + vc[vo.dsig() + l] * sa3[s * 4 * EL + EL + (size_t)l * E + e];
    }
    sa2[s * 3 * E + E + e] = sv;
    sa2[s * 3 * E + 2 * E + e] = sd;
  }
}

// Finiteness of the spectral state after the last step.
template <class T>
__global__ void k_check_spec(const cx<T> *sa3, size_t n3, const cx<T> *sa2, size_t n2, Ctl *ctl) {
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  T acc = 0;
  for (size_t q = idx; q < n3; q += stride) acc += sa3[q].x + sa3[q].y;
  for (size_t q = idx; q < n2; q += stride) acc += sa2[q].x + sa2[q].y;
  if (!isfinite(acc)) atomicMin(&ctl->bad_step, ctl->step);
}

// ---------------------------------------------------------------------------
// Layout conversions.
// ---------------------------------------------------------------------------
// reference grid (B, J, I, K) float64 -> internal [b][j][r0+k][I] of T,
// optionally times rcos[j] (vector analysis pre-scale).  K may be 1 for a
// surface field given as (B, J, I).
template <class T>
__global__ void k_pack_grid(const double *__restrict__ src, T *__restrict__ dst, int B, int J,
                            int I, int K, int R, int r0, const double *__restrict__ rcos) {
  const size_t total = (size_t)B * J * I * K;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (size_t)gridDim.x * blockDim.x) {
    // q enumerates the destination order (b, j, k, i) for coalesced stores
    const int i = (int)(q % I);
    size_t rest = q / I;
    const int k = (int)(rest % K);
    rest /= K;
    const int j = (int)(rest % J);
    const int b = (int)(rest / J);
    double v = src[(((size_t)b * J + j) * I + i) * K + k];
    if (rcos) v *= rcos[j];
    dst[(((size_t)b * J + j) * R + r0 + k) * I + i] = (T)v;
  }
}

// internal [b][j][r0+k][I] -> reference (B, J, I, K) float64.
// mode 0: copy; 1: ps' = ps_bar (exp(x) - 1) with ps_bar from psbar[b?][j][i];
// 2: ps' = ps_bar x (the linearised pair).
template <class T>
__global__ void k_unpack_grid(const T *__restrict__ src, double *__restrict__ dst, int B, int J,
                              int I, int K, int R, int r0, int mode, const T *psbar,
                              int psbar_rows, int psbar_row, int nbasic) {
  const size_t total = (size_t)B * J * I * K;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(q % K);
    size_t rest = q / K;
    const int i = (int)(rest % I);
    rest /= I;
    const int j = (int)(rest % J);
    const int b = (int)(rest / J);
    T v = src[(((size_t)b * J + j) * R + r0 + k) * I + i];
    if (mode != 0) {
      const int bb = nbasic > 1 ? b : 0;
      const T pb = psbar[(((size_t)bb * J + j) * psbar_rows + psbar_row) * I + i];
      v = mode == 1 ? pb * expm1(v) : pb * v;
    }
    dst[q] = (double)v;
  }
}

// reference spectral (B, M, M, K) complex128 -> packed [s][f][ncols] with
// column b*K + k (f fixed by the caller through the destination offset).
template <class T>
__global__ void k_pack_spec(const double *__restrict__ src, cx<T> *__restrict__ dst, int B, int M,
                            int K, int nf, int f, int ncols) {
  const int nspec = M * (M + 1) / 2;
  const size_t total = (size_t)nspec * B * K;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (size_t)gridDim.x * blockDim.x) {
    const int col = (int)(q % ((size_t)B * K));
    const int s = (int)(q / ((size_t)B * K));
    const int k = col / B, b = col - k * B;  // columns (level, member), member fastest
    int m = 0;
    while (spec_off(m + 1, M) <= s) ++m;
    const int n = m + (s - spec_off(m, M));
    const size_t si = ((((size_t)b * M + m) * M + n) * K + k) * 2;
    dst[((size_t)s * nf + f) * ncols + col] = mk((T)src[si], (T)src[si + 1]);
  }
}

template <class T>
__global__ void k_unpack_spec(const cx<T> *__restrict__ src, double *__restrict__ dst, int B,
                              int M, int K, int nf, int f, int ncols) {
  const size_t total = (size_t)B * M * M * K;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(q % K);
    size_t rest = q / K;
    const int n = (int)(rest % M);
    rest /= M;
    const int m = (int)(rest % M);
    const int b = (int)(rest / M);
    double re = 0, im = 0;
    if (n >= m) {
      const int s = spec_off(m, M) + n - m;
      const cx<T> v = src[((size_t)s * nf + f) * ncols + (size_t)k * B + b];
      re = (double)v.x;
      im = (double)v.y;
    }
    dst[2 * q] = re;
    dst[2 * q + 1] = im;
  }
}

// ps' (B, J, I) float64 -> pi' = log1p(ps'/ps_bar) (linear: ps'/ps_bar) into
// an internal grid row.
template <class T>
__global__ void k_pack_pi(const double *__restrict__ ps, T *__restrict__ dst, int B, int J, int I,
                          int R, int r0, const double *__restrict__ psbar, int nbasic, int linear) {
  const size_t total = (size_t)B * J * I;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % I);
    const int j = (int)((q / I) % J);
    const int b = (int)(q / ((size_t)I * J));
    const int bb = nbasic > 1 ? b : 0;
    const double pb = psbar[((size_t)bb * J + j) * I + i];
    const double v = ps ? ps[q] : 0.0;
    dst[(((size_t)b * J + j) * R + r0) * I + i] = (T)(linear ? v / pb : log1p(v / pb));
  }
}

// The deferred ps' update of the last step applied now (the same
// arithmetic as column_surface, so the next step sees bit-identical values):
// mode 1 assign (start-up), 2 Robert-Asselin filter, from the synthesised
// pi row r0 of g.  The caller clears pend afterwards.
template <class T>
__global__ void k_settle_ps(const T *__restrict__ g, int R, int r0, const T *__restrict__ bsg,
                            int brows, int brow, T *__restrict__ psa, T *__restrict__ psb,
                            const int *__restrict__ pend, T robert, int B, int J, int I, int nbasic) {
  const size_t tot = (size_t)B * J * I;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot;
       q += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % I);
    const size_t bj = q / I;
    const int j = (int)(bj % J), b = (int)(bj / J);
    const int mode = pend[b];
    if (mode == 0) continue;
    const int bb = nbasic > 1 ? b : 0;
    const T pb = bsg[(((size_t)bb * J + j) * brows + brow) * I + i];
    const T psn = pb * expm1(g[(bj * R + r0) * I + i]);
    const T pa = psa[q];
    psb[q] = (mode == 1) ? pa : pa + robert * (psb[q] - (T)2 * pa + psn);
    psa[q] = psn;
  }
}

// psa = psb = ps_bar (exp(pi) - 1) from a synthesised pi row (state.py:81-83)
template <class T>
__global__ void k_ps_from_pi(const T *__restrict__ g, int R, int r0, const T *__restrict__ bsg,
                             int brows, int brow, T *__restrict__ psa, T *__restrict__ psb, int B,
                             int J, int I, int nbasic) {
  const size_t tot = (size_t)B * J * I;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot;
       q += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % I);
    const size_t bj = q / I;
    const int j = (int)(bj % J), b = (int)(bj / J);
    const int bb = nbasic > 1 ? b : 0;
    const T pb = bsg[(((size_t)bb * J + j) * brows + brow) * I + i];
    const T v = pb * expm1(g[(bj * R + r0) * I + i]);
    psa[q] = v;
    psb[q] = v;
  }
}

// ---------------------------------------------------------------------------
// Basic-state terms (dynamics.py:54-93), FP64.  gsyn rows: vor (L), div (L);
// ggrad rows: d/dx ln ps, d/dy ln ps.  Writes the BRow block.
// ---------------------------------------------------------------------------
static __global__ __launch_bounds__(128) void k_basic_terms(
    const double *__restrict__ u, const double *__restrict__ v, const double *__restrict__ t,
    const double *__restrict__ q, const double *__restrict__ ps,  // reference layouts
    const double *__restrict__ gsyn, const double *__restrict__ ggrad, const double *__restrict__ vc,
    const double *__restrict__ mu, double *__restrict__ bsg, int J, int I, int L) {
  const int nch = I / (int)blockDim.x;  // blocks of 128 or 64 longitudes
  const int bj = blockIdx.x / nch;
  const int i = (blockIdx.x - bj * nch) * (int)blockDim.x + threadIdx.x;
  const int b = bj / J, j = bj - b * J;
  const VOff vo{L};
  const BRow br{L};
  const size_t col = ((size_t)b * J + j) * I + i;  // (b, j, i) in 2-D reference layout
  const double *vs = gsyn + ((size_t)(b * J + j) * 2 * L) * I + i;
  const double *gg = ggrad + ((size_t)(b * J + j) * 2) * I + i;
  double *o = bsg + ((size_t)(b * J + j) * br.rows()) * I + i;
  const double pix = gg[0], piy = gg[(size_t)I];
  const double f = 2.0 * OMEGA * mu[j];
  double S = 0;
  for (int k = L - 1; k >= 0; --k) {  // column sum from the top
    const size_t g = col * L + k;
    const double d = vs[(size_t)(L + k) * I];
    const double padv = u[g] * pix + v[g] * piy;
    S += vc[vo.dsig() + k] * (d + padv);
  }
  const double S0 = S;
  // bottom-up: S_k = S_0 - sum_{l<k} dsig (D + A); sdot_k = sigm_k S_0 - S_k
  double Sk = S0, su = 0, sv = 0;
  for (int k = 0; k < L; ++k) {
    const size_t g = col * L + k;
    const double d = vs[(size_t)(L + k) * I];
    const double uu = u[g], vv = v[g], tt = t[g], qq = q[g];
    const double padv = uu * pix + vv * piy;
    const double sdk = (k == 0) ? 0.0 : vc[vo.sigm() + k] * S0 - Sk;
    const double Sk1 = Sk - vc[vo.dsig() + k] * (d + padv);
    o[(size_t)(br.u() + k) * I] = uu;
    o[(size_t)(br.v() + k) * I] = vv;
    o[(size_t)(br.t() + k) * I] = tt;
    o[(size_t)(br.q() + k) * I] = qq;
    o[(size_t)(br.tv() + k) * I] = tt * (1.0 + EPSVT * qq);
    o[(size_t)(br.absvor() + k) * I] = vs[(size_t)k * I] + f;
    o[(size_t)(br.div() + k) * I] = d;
    o[(size_t)(br.x5() + k) * I] =
        vc[vo.w3() + k] * (k + 1 == L ? 0.0 : Sk1) + vc[vo.alpha() + k] * d + vc[vo.wP() + k] * padv;
    o[(size_t)(br.sdot() + k) * I] = sdk;
    su += vc[vo.dsig() + k] * uu;
    sv += vc[vo.dsig() + k] * vv;
    Sk = Sk1;
  }
  o[(size_t)br.sbu() * I] = su;
  o[(size_t)br.sbv() * I] = sv;
  o[(size_t)(br.sdot() + L) * I] = 0.0;
  o[(size_t)br.pix() * I] = pix;
  o[(size_t)br.piy() * I] = piy;
  o[(size_t)br.psbar() * I] = ps[col];
}

// Plain grid mean of basic Tv per level (vertical.py:206-216), fixed order.
static __global__ __launch_bounds__(256) void k_tbar(const double *__restrict__ bsg, double *__restrict__ tbar,
                                              int J, int I, int L) {
  __shared__ double red[256];
  const int b = blockIdx.x / L, k = blockIdx.x - b * L;
  const BRow br{L};
  double acc = 0;
  for (int p = threadIdx.x; p < J * I; p += 256) {
    const int j = p / I, i = p - j * I;
    acc += bsg[((size_t)(b * J + j) * br.rows() + br.tv() + k) * I + i];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) tbar[(size_t)b * L + k] = red[0] / (double)(J * I);
}

// g[row r0] <- log(g[row r0]) for every (b, j, i)  (ln ps of dynamics.py:74)
template <class T>
__global__ void k_log_row(T *g, int B, int J, int I, int R, int r0) {
  const size_t tot = (size_t)B * J * I;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot;
       q += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % I);
    const size_t bj = q / I;
    T *x = g + (bj * R + r0) * I + i;
    *x = log(*x);
  }
}

// zero K rows r0..r0+K-1 of an internal grid [b][j][R][I]
template <class T>
__global__ void k_zero_rows(T *g, int B, int J, int I, int R, int r0, int K) {
  const size_t tot = (size_t)B * J * K * I;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot;
       q += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % I);
    size_t rest = q / I;
    const int k = (int)(rest % K);
    const size_t bj = rest / K;
    g[(bj * R + r0 + k) * I + i] = (T)0;
  }
}

template <class S, class D>
__global__ void k_convert(const S *__restrict__ src, D *__restrict__ dst, size_t n) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x)
    dst[q] = (D)src[q];
}

}  // namespace lbm
