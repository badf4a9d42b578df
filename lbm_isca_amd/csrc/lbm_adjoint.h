// Adjoint of the linear time step: strictly linear tendencies and the
// linearised pi' <-> ps' pair, so the state is purely spectral and the
// Robert-Asselin filter acts on all five fields.  Every stage of the forward
// step is real-linear; each kernel here is its exact transpose for the
// Euclidean inner product of the stored reals (re/im of spectral and Fourier
// coefficients, grid values):
//   spectral step      -> k_step_adj  (M^T, V^T, tmtx^T, WW^T, RA^T)
//   Legendre analysis  -> k_syn_mfma with the analysis tables (gw P, gw H)
//   forward FFT (/N)   -> inverse FFT of X_m (1 or 1/2)/N       (mode 1)
//   grid tendencies    -> k_tend_adj  (levels top-down, suffix sums)
//   inverse FFT        -> forward FFT times N (1 or 2), no 1/N   (mode 2)
//   Legendre synthesis -> k_ana_mfma with the synthesis tables (P, H)
// (i m)^T = -(i m); the FFT factors are applied by the FFT kernels (their
// mode argument) and the 1/(a cos lat) factors of the vector fields by the
// grid kernels.
#pragma once
#include "lbm_mfma.h"
#include "lbm_model.h"

namespace lbm {

// ---------------------------------------------------------------------------
// Synthesis-kernel recipes (run with the analysis tables gw P, gw H and the
// per-degree table n(n+1)/a^2).
// ---------------------------------------------------------------------------
// transpose of the tendency analysis (AnaTendVD, AnaFluxDiv twice):
// (vor, div, t, q)* -> Fourier (fu' fv' ke xu' xv' gt yu' yv' g)*
struct AdjTend {
  static constexpr int NIN = 4, NOUT = 9;
  static constexpr bool hasA(int o) { return o != 4 && o != 7; }
  static constexpr bool hasB(int o) { return o == 0 || o == 1 || o == 4 || o == 7; }
  static constexpr bool scaled(int) { return false; }
  template <class T>
  __device__ __forceinline__ static void terms(const cx<T> *x, T nn1a2, T fm, cx<T> *A, cx<T> *B) {
    const cx<T> zv = x[0], zd = x[1], zt = x[2], zq = x[3];
    A[0] = im_mul(-fm, zd);  // fu': div = i m AP[fu'] + ...
    B[0] = zv;               //      vor = AH[fu'] + ...
    A[1] = im_mul(-fm, zv);  // fv': vor = i m AP[fv'] + ...
    B[1] = -zd;              //      div = -AH[fv'] + ...
    A[2] = nn1a2 * zd;       // ke
    A[3] = im_mul(fm, zt);   // xu': t = -i m AP[xu'] + ...
    B[4] = zt;               // xv'
    A[5] = zt;               // gt
    A[6] = im_mul(fm, zq);   // yu'
    B[7] = zq;               // yv'
    A[8] = zq;               // g
  }
};
// transpose of the surface analysis of gt_pi
struct AdjScalarFft {
  static constexpr int NIN = 1, NOUT = 1;
  static constexpr bool hasA(int) { return true; }
  static constexpr bool hasB(int) { return false; }
  static constexpr bool scaled(int) { return false; }
  template <class T>
  __device__ __forceinline__ static void terms(const cx<T> *x, T, T, cx<T> *A, cx<T> *) {
    A[0] = x[0];
  }
};

// ---------------------------------------------------------------------------
// Analysis-kernel recipes (run with the synthesis tables P, H and the
// per-degree table -a^2/(n(n+1)), 0 at n = 0).
// ---------------------------------------------------------------------------
// transpose of SynLevel: Fourier (U V vor div T Q)* -> (vor, div, t, q)*,
// added to the state adjoint
struct AdjSynLevel {
  static constexpr int NIN = 6, NOUT = 4;
  static constexpr bool ACC = true;
  static constexpr bool needP(int) { return true; }
  static constexpr bool needH(int f) { return f < 2; }
  template <class T>
  __device__ __forceinline__ static void combine(const cx<T> *AP, const cx<T> *AH, T fm, T il,
                                                 cx<T> *o) {
    // U = P (i m il div) - H (il vor),  V = P (i m il vor) + H (il div)
    o[0] = (-il) * AH[0] + il * im_mul(-fm, AP[1]) + AP[2];
    o[1] = il * im_mul(-fm, AP[0]) + il * AH[1] + AP[3];
    o[2] = AP[4];
    o[3] = AP[5];
  }
};
// transpose of SynSurf: (d/dx pi, d/dy pi, pi, sum u, sum v, sum D)* ->
// (pi, sum vor, sum div)*
struct AdjSynSurf {
  static constexpr int NIN = 6, NOUT = 3;
  static constexpr bool ACC = false;
  static constexpr bool needP(int f) { return f != 1; }
  static constexpr bool needH(int f) { return f == 1 || f == 3 || f == 4; }
  template <class T>
  __device__ __forceinline__ static void combine(const cx<T> *AP, const cx<T> *AH, T fm, T il,
                                                 cx<T> *o) {
    o[0] = im_mul(-fm, AP[0]) + AH[1] + AP[2];
    o[1] = (-il) * AH[3] + il * im_mul(-fm, AP[4]);
    o[2] = il * im_mul(-fm, AP[3]) + il * AH[4] + AP[5];
  }
};
// transpose of vorticity -> streamfunction on the grid (inverse Laplacian,
// then wave_to_grid): grid weights -> the linear functional of vorticity
struct AdjPsi {
  static constexpr int NIN = 1, NOUT = 1;
  static constexpr bool ACC = false;
  static constexpr bool needP(int) { return true; }
  static constexpr bool needH(int) { return false; }
  template <class T>
  __device__ __forceinline__ static void combine(const cx<T> *AP, const cx<T> *, T, T il,
                                                 cx<T> *o) {
    o[0] = il * AP[0];
  }
};

// Transposes of the API transforms (spectral.py:210-276 forward).
// uv_to_vordiv^T (synthesis kernel, analysis tables): (vor*, div*) ->
// Fourier (U*, V*), times 1/(a cos lat) for the forward's pre-scale
struct AdjVorDiv {
  static constexpr int NIN = 2, NOUT = 2;
  static constexpr bool hasA(int) { return true; }
  static constexpr bool hasB(int) { return true; }
  static constexpr bool scaled(int) { return true; }
  template <class T>
  __device__ __forceinline__ static void terms(const cx<T> *x, T, T fm, cx<T> *A, cx<T> *B) {
    A[0] = im_mul(-fm, x[1]);  // vor = i m AP[V] + AH[U], div = i m AP[U] - AH[V]
    B[0] = x[0];
    A[1] = im_mul(-fm, x[0]);
    B[1] = -x[1];
  }
};
// psichi_to_uv^T (analysis kernel, synthesis tables): (U*, V*) pre-scaled
// by 1/(a cos lat) -> (psi*, chi*)
struct AdjPsiChi {
  static constexpr int NIN = 2, NOUT = 2;
  static constexpr bool ACC = false;
  static constexpr bool needP(int) { return true; }
  static constexpr bool needH(int) { return true; }
  template <class T>
  __device__ __forceinline__ static void combine(const cx<T> *AP, const cx<T> *AH, T fm, T,
                                                 cx<T> *o) {
    // U = P (i m chi) - H psi,  V = P (i m psi) + H chi
    o[0] = im_mul(-fm, AP[1]) - AH[0];
    o[1] = im_mul(-fm, AP[0]) + AH[1];
  }
};
// hgrad^T (analysis kernel, synthesis tables): (gx*, gy*) pre-scaled by
// 1/(a cos lat) -> w*
struct AdjGrad {
  static constexpr int NIN = 2, NOUT = 1;
  static constexpr bool ACC = false;
  static constexpr bool needP(int f) { return f == 0; }
  static constexpr bool needH(int f) { return f == 1; }
  template <class T>
  __device__ __forceinline__ static void combine(const cx<T> *AP, const cx<T> *AH, T fm, T,
                                                 cx<T> *o) {
    o[0] = im_mul(-fm, AP[0]) + AH[1];  // gx = P (i m w), gy = H w
  }
};

// ---------------------------------------------------------------------------
// Transposed grid tendencies of the linear model.  gout* rows: fu' fv' ke
// xu' xv' gt yu' yv' g (L each), gt_pi, (pi_prev, not read).  gin* rows: U V
// vor div T Q (L each), d/dx pi, d/dy pi, pi, sum u, sum v, sum D, with the
// 1/(a cos lat) factor of the scaled synthesis outputs applied.  The forcing
// adjoint accumulates gt* ([e][j][L][I]).
// ---------------------------------------------------------------------------
template <class T>
struct TendAdjArgs {
  const T *gout_a;  // [e][j][9L+2][I]
  T *gin_a;         // [e][j][6L+6][I]
  T *frc_a;         // [e][j][L][I], accumulated
  const T *bsg;     // [eb][j][9L+6][I]
  const T *vc, *tref, *rcos;
  int E, J, I, L, nbasic;
  ActiveMask am;
  // the default model about a stored trajectory (null: the tangent-linear
  // model about rest): this step's trajectory [e][j][2L+2][I] (u, v per
  // level, ps' of both levels after the deferred update), the carried grid
  // adjoints of ps' at both levels, and what the mode of the deferred
  // update needs
  const T *traj;
  T *psa_a, *psb_a;
  const int *nsteps;
  int inistp, step, full_ke;
  T robert;
};

template <class T>
__global__ __launch_bounds__(128) void k_tend_adj(TendAdjArgs<T> a) {
  const int nch = a.I / (int)blockDim.x;  // blocks of 128 or 64 longitudes
  const int ej = blockIdx.x / nch;
  const int i = (blockIdx.x - ej * nch) * (int)blockDim.x + threadIdx.x;
  const int e = ej / a.J, j = ej - e * a.J;
  if (!a.am(e)) return;  // a finished member contributes nothing
  const int L = a.L, I = a.I, J = a.J;
  const int eb = a.nbasic > 1 ? e : 0;
  const VOff vo{L};
  const BRow br{L};
  const T *vc = a.vc;
  const T *tref = a.tref + (size_t)eb * L;
  const T *bs = a.bsg + (size_t)(eb * J + j) * br.rows() * I + i;
  const T *go = a.gout_a + (size_t)(e * J + j) * (9 * L + 2) * I + i;
  T *gi = a.gin_a + (size_t)(e * J + j) * (6 * L + 6) * I + i;
  T *fg = a.frc_a + (size_t)(e * J + j) * L * I + i;
  auto B = [&](int row) { return bs[(size_t)row * I]; };
  auto O = [&](int f, int k) { return go[(size_t)(f * L + k) * I]; };
  const T rc = a.rcos[j];
  const T eps = (T)EPSVT, kap = (T)KAPPA, rgas = (T)R_DRY, h = (T)0.5;
  const T bpix = B(br.pix()), bpiy = B(br.piy());
  const T *trj = a.traj ? a.traj + (size_t)(e * J + j) * (2 * L + 2) * I + i : nullptr;

  // adjoints of the four tendencies carrying vertical advection (u, v, T, q)
  auto W = [&](int k, T *w) {
    w[0] = O(0, k) * rc;
    w[1] = O(1, k) * rc;
    w[2] = O(5, k);
    w[3] = O(8, k);
  };
  T Pst = 0, PDst = 0, SKE = 0, S0st = 0, D0st = 0, pixst = 0, piyst = 0;
  T wk[4], wn[4] = {0, 0, 0, 0}, wp[4] = {0, 0, 0, 0};
  // the basic state and tendency adjoints of level k-1 are loaded while
  // level k is formed; level k+1's basic u, v, T, q and sigma-dot come from
  // the previous iteration (levels run top-down)
  struct ALev {
    T u = 0, v = 0, t = 0, q = 0, tv = 0, av = 0, dv = 0, x5 = 0, sdlo = 0;
    T ke = 0, xu = 0, xv = 0, yu = 0, yv = 0;
  };
  auto alev = [&](int k) {
    ALev b;
    b.u = B(br.u() + k);
    b.v = B(br.v() + k);
    b.t = B(br.t() + k);
    b.q = B(br.q() + k);
    b.tv = B(br.tv() + k);
    b.av = B(br.absvor() + k);
    b.dv = B(br.div() + k);
    b.x5 = B(br.x5() + k);
    b.sdlo = B(br.sdot() + k);
    b.ke = O(2, k);
    b.xu = O(3, k) * rc;
    b.xv = O(4, k) * rc;
    b.yu = O(6, k) * rc;
    b.yv = O(7, k) * rc;
    return b;
  };
  W(L - 1, wk);
  ALev cur = alev(L - 1);
  T bsd_hi = B(br.sdot() + L);
  T bun = 0, bvn = 0, btn = 0, bqn = 0;
  for (int k = L - 1; k >= 0; --k) {
    if (k > 0) W(k - 1, wp);
    const ALev nxt = k > 0 ? alev(k - 1) : ALev{};
    const T ds = vc[vo.dsig() + k], rds = (T)1 / ds;
    const T alpha = vc[vo.alpha() + k], w3 = vc[vo.w3() + k], wP = vc[vo.wP() + k];
    const T rcpg = rgas * vc[vo.cpg() + k], dlog3 = vc[vo.dlog3() + k];
    const T tr = tref[k];
    const T bu = cur.u, bv = cur.v, bt = cur.t, bq = cur.q;
    const T btv = cur.tv, babsv = cur.av, bdiv = cur.dv;
    const T bx5 = cur.x5;
    const T bsd_lo = cur.sdlo;
    const bool top = (k + 1 == L);
    const T fu_a = wk[0], fv_a = wk[1], gt_a = wk[2], g_a = wk[3];
    const T ke_a = cur.ke, xu_a = cur.xu, xv_a = cur.xv;
    const T yu_a = cur.yu, yv_a = cur.yv;
    T up_a = 0, vp_a = 0, vor_a = 0, d_a = 0, tp_a = 0, qp_a = 0, tv_a = 0, padv_a = 0;
    const T tdev = bt - tr, tvdev = btv - tr;
    // temperature: adiabatic term -kappa (tv x5 + tv_bar x5'), energy
    // conversion kappa Tr (alpha D + w3 (D_0 - PD_{k+1})), advection
    const T x5p_a = -kap * btv * gt_a;
    tv_a += -kap * bx5 * gt_a;
    d_a += kap * tr * alpha * gt_a;
    D0st += kap * tr * w3 * gt_a;
    const T PD_a = PDst - kap * tr * w3 * gt_a;  // adjoint of PD_{k+1}
    tp_a += bdiv * gt_a;
    d_a += tdev * gt_a;
    d_a += alpha * x5p_a;
    padv_a += wP * x5p_a;
    const T S_a = w3 * x5p_a;  // adjoint of S_{k+1} = S_0 - P_{k+1}
    up_a += tdev * xu_a;
    tp_a += bu * xu_a;
    vp_a += tdev * xv_a;
    tp_a += bv * xv_a;
    // momentum
    vp_a += babsv * fu_a;
    vor_a += bv * fu_a;
    tv_a += -rcpg * bpix * fu_a;
    pixst += -rcpg * tvdev * fu_a;
    up_a += -babsv * fv_a;
    vor_a += -bu * fv_a;
    tv_a += -rcpg * bpiy * fv_a;
    piyst += -rcpg * tvdev * fv_a;
    // kinetic energy u_bar u' + v_bar v' (+ u' u'/2 + v' v'/2 of the
    // default model: about the trajectory, u_traj u' + v_traj v') and the
    // virtual geopotential
    up_a += bu * ke_a;
    vp_a += bv * ke_a;
    if (trj && a.full_ke) {
      up_a += trj[(size_t)k * I] * ke_a;
      vp_a += trj[(size_t)(L + k) * I] * ke_a;
    }
    const T x_a = rgas * alpha * ke_a + rgas * dlog3 * SKE;
    SKE += ke_a;
    tv_a += x_a;
    tp_a -= x_a;
    // humidity
    up_a += bq * yu_a;
    qp_a += bu * yu_a;
    vp_a += bq * yv_a;
    qp_a += bv * yv_a;
    qp_a += bdiv * g_a;
    d_a += bq * g_a;
    // virtual temperature tv' = t' + eps (t' q_bar + t_bar q')
    tp_a += ((T)1 + eps * bq) * tv_a;
    qp_a += eps * bt * tv_a;
    // vertical advection of the perturbation by the basic sigma-dot: level
    // k reads f_{k-1}, f_k, f_{k+1}; the bottom and top fluxes are zero
    {
      T *fa[4] = {&up_a, &vp_a, &tp_a, &qp_a};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        T acc = -wk[q] * (bsd_hi - bsd_lo) * rds;
        if (!top) acc += h * bsd_hi * (wk[q] * rds - wn[q] / vc[vo.dsig() + k + 1]);
        if (k > 0) acc += h * bsd_lo * (wp[q] / vc[vo.dsig() + k - 1] - wk[q] * rds);
        *fa[q] += acc;
      }
    }
    // perturbation sigma-dot at half level k+1 advects the basic fields of
    // levels k and k+1
    T sd_a = 0;
    if (!top) {
      const T r1 = (T)1 / vc[vo.dsig() + k + 1];
      sd_a = h * ((bun - bu) * (wk[0] * rds + wn[0] * r1) + (bvn - bv) * (wk[1] * rds + wn[1] * r1) +
                  (btn - bt) * (wk[2] * rds + wn[2] * r1) + (bqn - bq) * (wk[3] * rds + wn[3] * r1));
      S0st += -((T)1 - vc[vo.sigm() + k + 1]) * sd_a;
    }
    // P_{k+1} = P_k + dsig (D + A'), PD_{k+1} = PD_k + dsig D
    const T P_a = Pst + sd_a - S_a;
    S0st += S_a;
    d_a += ds * (P_a + PD_a);
    padv_a += ds * P_a;
    Pst = P_a;
    PDst = PD_a;
    // A' = u' d/dx ln ps_bar + u_bar d/dx pi' + v' d/dy ln ps_bar + v_bar d/dy pi'
    up_a += bpix * padv_a;
    pixst += bu * padv_a;
    vp_a += bpiy * padv_a;
    piyst += bv * padv_a;

    gi[(size_t)(0 * L + k) * I] = up_a * rc;
    gi[(size_t)(1 * L + k) * I] = vp_a * rc;
    gi[(size_t)(2 * L + k) * I] = vor_a;
    gi[(size_t)(3 * L + k) * I] = d_a;
    gi[(size_t)(4 * L + k) * I] = tp_a;
    gi[(size_t)(5 * L + k) * I] = qp_a;
    fg[(size_t)k * I] += gt_a;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      wn[q] = wk[q];
      wk[q] = wp[q];
    }
    bun = bu;
    bvn = bv;
    btn = bt;
    bqn = bq;
    bsd_hi = bsd_lo;
    cur = nxt;
  }
  // gt_pi = D_0 - S_0;  S_0 = sum D + bpix sum u + bpiy sum v + pix sbu + piy sbv
  const T gpi_a = go[(size_t)(9 * L) * I];
  D0st += gpi_a;
  S0st -= gpi_a;
  pixst += B(br.sbu()) * S0st;
  piyst += B(br.sbv()) * S0st;
  // the default model: transposes of pi_prev = log1p(ps'_b / ps_bar) and of
  // the deferred update at the top of this step (ps'_a = ps_bar expm1(pi),
  // ps'_b from the start-up assignment or the Robert-Asselin filter of the
  // previous levels), carrying the grid adjoints of ps' to the step before
  T pig_a = 0;
  if (trj) {
    const size_t pidx = (size_t)(e * J + j) * I + i;
    const T psbar = B(br.psbar());
    const T ca = psbar + trj[(size_t)(2 * L) * I], cb = psbar + trj[(size_t)(2 * L + 1) * I];
    const T psa = a.psa_a[pidx];
    const T psb = a.psb_a[pidx] + go[(size_t)(9 * L + 1) * I] / cb;
    const int k = a.step;
    const bool upd = k >= 1 && (k - 1) < a.inistp + a.nsteps[e];
    const int mode = upd ? ((k - 1) < a.inistp ? 1 : 2) : 0;
    if (mode == 2) {
      pig_a = ca * (psa + a.robert * psb);
      a.psa_a[pidx] = ((T)1 - (T)2 * a.robert) * psb;
      a.psb_a[pidx] = a.robert * psb;
    } else if (mode == 1) {
      pig_a = ca * psa;
      a.psa_a[pidx] = psb;
      a.psb_a[pidx] = 0;
    } else {
      a.psb_a[pidx] = psb;
    }
  }
  gi[(size_t)(6 * L) * I] = pixst * rc;
  gi[(size_t)(6 * L + 1) * I] = piyst * rc;
  gi[(size_t)(6 * L + 2) * I] = pig_a;
  gi[(size_t)(6 * L + 3) * I] = bpix * S0st * rc;
  gi[(size_t)(6 * L + 4) * I] = bpiy * S0st * rc;
  gi[(size_t)(6 * L + 5) * I] = S0st + D0st;
}

// ---------------------------------------------------------------------------
// Transposed spectral step.  On entry aa3/aa2 and ab3/ab2 hold the adjoints
// of the state after the step (Sa_new, Sb_new); on exit those of the state
// before it.  Also folded in: the synthesis adjoint of the column sums and
// pi from the step after this one (s3a: pi*, (sum dsig vor)*, (sum dsig
// div)*), and the response weight of the state after this step.  The
// tendency adjoints go to at3 ([s][4][L][E]) and at2 ([s][2][E], field 0).
// ---------------------------------------------------------------------------
template <class T>
struct StepAdjArgs {
  cx<T> *aa3, *ab3;  // [s][4][L][E]
  cx<T> *aa2, *ab2;  // [s][E] pi
  const cx<T> *s3a;  // [s][3][E]
  cx<T> *at3, *at2;
  const T *kv, *kd, *ktt, *kq, *minv, *vinv, *vc, *tref, *lap;
  const short *n_of_s;
  const cx<T> *obs;  // [nspec] response functional of the vorticity, or null
  const T *wlev;     // [L] response weight per level
  const T *wstep;    // [nwstep] response weight per step
  int nwstep;
  const Ctl *ctl;
  const int *nsteps;
  int E, M, nspec, nbasic, inistp, vmask, skip_low_n;
  int L;     // levels
  int ecnt;  // members covered by the launch (<= E)
  T delt, robert;
  // 1: the default model (pi at t-dt is the analysed ps'_b: its adjoint goes
  // to the tendency field 1, and pi has no spectral Robert-Asselin filter)
  int traj;
};

template <class T, int LT, bool EXACT, class P>
__device__ inline void matvec_t(P A, const cx<T> *x, cx<T> *y, int L) {
#pragma unroll
  for (int c = 0; c < LT; ++c) if (EXACT || c < L) {
    cx<T> acc = zero<T>();
#pragma unroll
    for (int r = 0; r < LT; ++r) if (EXACT || r < L) fma_acc(acc, A[r * L + c], x[r]);
    y[c] = acc;
  }
}

// LT, EXACT, NG (lane mapping): as in k_step.
template <class T, int LT, bool EXACT = true, int NG = 0>
__global__ __launch_bounds__(64) void k_step_adj(StepAdjArgs<T> a) {
  const int L = EXACT ? LT : a.L;
  const int E = a.E;
  int s, e, ng_n = 0;
  if constexpr (NG > 0) {
    if (!ng_lane<NG>(a.M, a.ecnt, s, e, ng_n)) return;
  } else {
    const int idx = blockIdx.x * 64 + threadIdx.x;
    s = idx / a.ecnt;
    e = idx - s * a.ecnt;
  }
  if (s >= a.nspec) return;
  const int st = a.ctl->step;
  const size_t EL = (size_t)E * L;
  // level arrays through buffer views based at the wave's lowest coefficient
  // (per-lane offset + wave-uniform level offset), as in k_step
  const int s_w = NG > 0 ? __builtin_amdgcn_readfirstlane(ng_n) : __builtin_amdgcn_readfirstlane(s);
  const size_t wbase = (size_t)s_w * 4 * EL;
  const unsigned voff = (unsigned)(((size_t)(s - s_w) * 4 * EL + e) * sizeof(cx<T>));
  const size_t wbytes = (size_t)(a.nspec - s_w) * 4 * EL * sizeof(cx<T>);
  const unsigned wrange = wbytes < 0xffffffffull ? (unsigned)wbytes : 0xffffffffu;
  const BufView bA(a.aa3 + wbase, wrange), bB(a.ab3 + wbase, wrange), bT(a.at3 + wbase, wrange);
  auto lo = [&](int f, int l) __attribute__((always_inline)) {
    return (unsigned)(((size_t)f * EL + (size_t)l * E) * sizeof(cx<T>));
  };
  const cx<T> *const tag = nullptr;  // selects the precision of BufView::ld
  const VOff vo{L};
  using RU = typename RoPtr<T, true>::type;  // wave-uniform: through the scalar cache
  const RU vc = (RU)a.vc;
  const size_t i2 = (size_t)s * E + e;
  // blockIdx.y: 0 -> pi, divergence, temperature (the coupled solve);
  // 1 -> vorticity and humidity (independent of it), as in k_step
  const int role = blockIdx.y;
  const cx<T> spi = a.s3a[(size_t)s * 3 * E + e];
  const cx<T> svor = a.s3a[(size_t)s * 3 * E + E + e];
  const cx<T> sdiv = a.s3a[(size_t)s * 3 * E + 2 * E + e];
  cx<T> *at2 = a.at2 + (size_t)s * 2 * E + e;
  const bool active = st < a.inistp + a.nsteps[e];
  if (!active) {
    // a member past its length: the step is the identity
    if (role == 1) {
#pragma unroll
      for (int l = 0; l < LT; ++l) if (EXACT || l < L) {
        bA.st(voff, lo(0, l), bA.ld(voff, lo(0, l), tag) + vc[vo.dsig() + l] * svor);
        bT.st(voff, lo(0, l), zero<T>());
        bT.st(voff, lo(3, l), zero<T>());
      }
      return;
    }
#pragma unroll
    for (int l = 0; l < LT; ++l) if (EXACT || l < L) {
      bA.st(voff, lo(1, l), bA.ld(voff, lo(1, l), tag) + vc[vo.dsig() + l] * sdiv);
      bT.st(voff, lo(1, l), zero<T>());
      bT.st(voff, lo(2, l), zero<T>());
    }
    a.aa2[i2] = a.aa2[i2] + spi;
    *at2 = zero<T>();
    if (a.traj) at2[E] = zero<T>();
    return;
  }
  const bool oinit = st < a.inistp;
  const T dt = oinit ? a.delt / (T)(1 << (a.inistp - st)) : a.delt;
  const int set = oinit ? st : a.inistp;
  const T dt2 = (T)2 * dt;
  const int n = NG > 0 ? __builtin_amdgcn_readfirstlane(ng_n) : a.n_of_s[s];
  const int eb = a.nbasic > 1 ? e : 0;
  const T *tref = a.tref + (size_t)eb * L;
  const T *ktt = a.ktt + (size_t)n * L, *kd = a.kd + (size_t)n * L;
  const T *kv = a.kv + (size_t)n * L, *kq = a.kq + (size_t)n * L;
  const T lapn = a.lap[n];
  const T rg = (T)R_DRY, kap = (T)KAPPA, gam = a.robert;
  const RU vinv = (RU)(a.vinv + (size_t)set * 4 * L * L);
  const bool vd_low = !(a.skip_low_n && n <= 1);
  const bool resp = a.obs != nullptr;
  const T wr = (resp && st < a.nwstep) ? a.wstep[st] : (T)0;
  const cx<T> ob = resp ? a.obs[s] : zero<T>();
  // Robert-Asselin with the start-up assignment:
  //   Sb_new = oinit ? Sa : (1 - 2g) Sa + g Sb + g S,  Sa_new = S
  const T ca = oinit ? (T)1 : (T)1 - (T)2 * gam;
  const T cb = oinit ? (T)0 : gam;
  const T cn = oinit ? (T)0 : gam;
  cx<T> x[LT], y[LT], dbar[LT];

  if (role == 1) {
  // vorticity: vor = Vv (vor_b + 2dt wt_vor) / (1 + 2dt kv)
#pragma unroll
  for (int l = 0; l < LT; ++l) if (EXACT || l < L)
    x[l] = bA.ld(voff, lo(0, l), tag) + vc[vo.dsig() + l] * svor + (resp ? wr * a.wlev[l] : (T)0) * ob +
           cn * bB.ld(voff, lo(0, l), tag);
  if ((a.vmask & 1) && vd_low) {
    matvec_t<T, LT, EXACT>(vinv, x, y, L);
  } else {
#pragma unroll
    for (int l = 0; l < LT; ++l) if (EXACT || l < L) y[l] = x[l];
  }
#pragma unroll
  for (int l = 0; l < LT; ++l) if (EXACT || l < L) {
    const cx<T> yl = ((T)1 / ((T)1 + dt2 * kv[l])) * y[l];
    const cx<T> bb = bB.ld(voff, lo(0, l), tag);
    bB.st(voff, lo(0, l), cb * bb + yl);
    bA.st(voff, lo(0, l), ca * bb);
    bT.st(voff, lo(0, l), dt2 * yl);
  }
  // humidity
#pragma unroll
  for (int l = 0; l < LT; ++l) if (EXACT || l < L) x[l] = bA.ld(voff, lo(3, l), tag) + cn * bB.ld(voff, lo(3, l), tag);
  if (a.vmask & 8) {
    matvec_t<T, LT, EXACT>(vinv + 3 * L * L, x, y, L);
  } else {
#pragma unroll
    for (int l = 0; l < LT; ++l) if (EXACT || l < L) y[l] = x[l];
  }
#pragma unroll
  for (int l = 0; l < LT; ++l) if (EXACT || l < L) {
    const cx<T> yl = ((T)1 / ((T)1 + dt2 * kq[l])) * y[l];
    const cx<T> bb = bB.ld(voff, lo(3, l), tag);
    bB.st(voff, lo(3, l), cb * bb + yl);
    bA.st(voff, lo(3, l), ca * bb);
    bT.st(voff, lo(3, l), dt2 * yl);
  }
  return;
  }
  // pi: pi = pi_b + 2dt (wt_pi - sum dsig dbar)
  const cx<T> bpi = a.traj ? zero<T>() : a.ab2[i2];
  const cx<T> pin_a = a.aa2[i2] + spi + cn * bpi;
  a.aa2[i2] = ca * bpi;
  cx<T> pib_a = cb * bpi + pin_a;
  // divergence: D = Vd (2 dbar - D_b)
#pragma unroll
  for (int l = 0; l < LT; ++l) if (EXACT || l < L) x[l] = bA.ld(voff, lo(1, l), tag) + vc[vo.dsig() + l] * sdiv + cn * bB.ld(voff, lo(1, l), tag);
  if ((a.vmask & 2) && vd_low) {
    matvec_t<T, LT, EXACT>(vinv + 1 * L * L, x, y, L);
  } else {
#pragma unroll
    for (int l = 0; l < LT; ++l) if (EXACT || l < L) y[l] = x[l];
  }
#pragma unroll
  for (int l = 0; l < LT; ++l) if (EXACT || l < L) {
    const cx<T> bb = bB.ld(voff, lo(1, l), tag);
    bA.st(voff, lo(1, l), ca * bb);
    bB.st(voff, lo(1, l), cb * bb - y[l]);
    dbar[l] = (T)2 * y[l];
  }
  // temperature: T = Vt (T_b + 2dt (wt_t - tmtx dbar)) / ft
#pragma unroll
  for (int l = 0; l < LT; ++l) if (EXACT || l < L) x[l] = bA.ld(voff, lo(2, l), tag) + cn * bB.ld(voff, lo(2, l), tag);
  if (a.vmask & 4) {
    matvec_t<T, LT, EXACT>(vinv + 2 * L * L, x, y, L);
  } else {
#pragma unroll
    for (int l = 0; l < LT; ++l) if (EXACT || l < L) y[l] = x[l];
  }
#pragma unroll
  for (int l = 0; l < LT; ++l) if (EXACT || l < L) {
    const T ft = (T)1 + dt2 * ktt[l];
    const cx<T> yl = ((T)1 / ft) * y[l];
    const cx<T> bb = bB.ld(voff, lo(2, l), tag);
    bA.st(voff, lo(2, l), ca * bb);
    bB.st(voff, lo(2, l), cb * bb + yl);
    y[l] = dt2 * yl;  // wt_t* so far, and -(tmtx dbar)*
    bT.st(voff, lo(2, l), y[l]);
  }
  // (tmtx dbar)_k = kappa Tr_k (alpha_k dbar_k + w3_k sum_{l>k} dsig_l dbar_l)
  {
    cx<T> pre = zero<T>();
#pragma unroll
    for (int l = 0; l < LT; ++l) if (EXACT || l < L) {
      const cx<T> tt = (kap * tref[l] * vc[vo.alpha() + l]) * y[l] + vc[vo.dsig() + l] * pre;
      pre = pre + (kap * tref[l] * vc[vo.w3() + l]) * y[l];
      dbar[l] = dbar[l] - tt;
    }
  }
  const cx<T> wtp = dt2 * pin_a;
#pragma unroll
  for (int l = 0; l < LT; ++l) if (EXACT || l < L) dbar[l] = dbar[l] - vc[vo.dsig() + l] * wtp;
  // dbar = M(n)^-1 rhs,  rhs = (1 + dt kd) ft D_b + dt (ft wt_d - lap phy),
  // phy_l = phi_l + ft_l R Tr_l cpg_l (pi_b + dt wt_pi)
  matvec_t<T, LT, EXACT>(a.minv + (((size_t)set * a.nbasic + eb) * a.M + n) * L * L, dbar, x, L);
  cx<T> wdpix_a = zero<T>();
#pragma unroll
  for (int l = 0; l < LT; ++l) if (EXACT || l < L) {
    const T ft = (T)1 + dt2 * ktt[l];
    bB.st(voff, lo(1, l), bB.ld(voff, lo(1, l), tag) + (((T)1 + dt * kd[l]) * ft) * x[l]);
    bT.st(voff, lo(1, l), (dt * ft) * x[l]);
    y[l] = (-dt * lapn) * x[l];
    wdpix_a = wdpix_a + (ft * (rg * tref[l] * vc[vo.cpg() + l])) * y[l];
  }
  pib_a = pib_a + wdpix_a;
  if (a.traj)
    at2[E] = pib_a;  // pi_prev*: to the grid adjoint of ps'_b (k_tend_adj)
  else
    a.ab2[i2] = pib_a;
  *at2 = wtp + dt * wdpix_a;
  // phi_l = sum_{k<l} R dlog3_k w_k + R alpha_l w_l,  w = (1 + dt ktt) T_b + dt wt_t
  {
    cx<T> suf = zero<T>();
#pragma unroll
    for (int l = LT - 1; l >= 0; --l) if (EXACT || l < L) {
      const cx<T> w_a = (rg * vc[vo.alpha() + l]) * y[l] + (rg * vc[vo.dlog3() + l]) * suf;
      suf = suf + y[l];
      bB.st(voff, lo(2, l), bB.ld(voff, lo(2, l), tag) + ((T)1 + dt * ktt[l]) * w_a);
      bT.st(voff, lo(2, l), bT.ld(voff, lo(2, l), tag) + dt * w_a);
    }
  }
}

// Adds the column-sum and pi adjoints s3a into the state adjoint:
// vor*_l += dsig_l (sum vor)*, div*_l += dsig_l (sum div)*, pi* += pi*.
template <class T>
__global__ void k_vsum_adj(const cx<T> *__restrict__ s3a, cx<T> *__restrict__ aa3,
                           cx<T> *__restrict__ aa2, const T *__restrict__ vc, int E, int L,
                           int nspec) {
  const size_t tot = (size_t)nspec * E;
  const VOff vo{L};
  const size_t EL = (size_t)E * L;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot;
       q += (size_t)gridDim.x * blockDim.x) {
    const size_t s = q / E, e = q - s * E;
    const cx<T> pa = s3a[s * 3 * E + e], zv = s3a[s * 3 * E + E + e];
    const cx<T> zd = s3a[s * 3 * E + 2 * E + e];
    aa2[s * E + e] = aa2[s * E + e] + pa;
    for (int l = 0; l < L; ++l) {
      const T ds = vc[vo.dsig() + l];
      cx<T> *pv = aa3 + s * 4 * EL + (size_t)l * E + e;
      pv[0] = pv[0] + ds * zv;
      pv[EL] = pv[EL] + ds * zd;
    }
  }
}

// Response J_e += w_step sum_l w_l <obs, vor_l> over (re, im) of every
// coefficient of the state after the step; one block per member, fixed
// summation order, accumulated in FP64.
template <class T>
__global__ __launch_bounds__(256) void k_response(const cx<T> *__restrict__ sa3,
                                                  const cx<T> *__restrict__ obs,
                                                  const T *__restrict__ wlev,
                                                  const T *__restrict__ wstep, int nwstep, int E,
                                                  int L, int nspec, const Ctl *ctl,
                                                  const int *nsteps, int inistp,
                                                  double *__restrict__ J) {
  __shared__ double red[256];
  const int e = blockIdx.x;
  const int st = ctl->step;
  const double w = (st >= 0 && st < nwstep) ? (double)wstep[st] : 0.0;
  const bool active = st < inistp + nsteps[e];
  double acc = 0;
  if (w != 0.0 && active) {
    const size_t EL = (size_t)E * L;
    for (int l = 0; l < L; ++l) {
      const double wl = (double)wlev[l];
      if (wl == 0.0) continue;
      double al = 0;
      for (int s = threadIdx.x; s < nspec; s += 256) {
        const cx<T> v = sa3[(size_t)s * 4 * EL + (size_t)l * E + e];
        al += (double)obs[s].x * (double)v.x + (double)obs[s].y * (double)v.y;
      }
      acc += wl * al;
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) J[e] += w * red[0];
}

// out = a + b over n complex values (the initial Sa and Sb are one state)
template <class T>
__global__ void k_add_cx(const cx<T> *__restrict__ a, const cx<T> *__restrict__ b,
                         cx<T> *__restrict__ out, size_t n) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x)
    out[q] = a[q] + b[q];
}

static __global__ void k_tick_back(Ctl *ctl) { ctl->step -= 1; }

}  // namespace lbm
