// Host side of the gfx950 time-stepping core: FP64 setup (Gaussian grid,
// associated Legendre tables, Simmons-Burridge operators, semi-implicit and
// vertical-diffusion inverses), device buffers, the C ABI of
// include/lbm_hip.h and the step loop (eager or hipGraph replay).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <climits>
#include <cstddef>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/lbm_hip.h"
#include "lbm_adjoint.h"
#include "lbm_history.h"
#include "lbm_levels.h"
#include "lbm_util.h"

using namespace lbm;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return fail(LBM_ERR_RUNTIME, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                  __FILE__, __LINE__);                                                 \
  } while (0)

#define CHK(expr)            \
  do {                       \
    int _r = (expr);         \
    if (_r != LBM_OK) return _r; \
  } while (0)

// ---------------------------------------------------------------------------
// FP64 host setup
// ---------------------------------------------------------------------------
// Gauss-Legendre nodes/weights on [-1,1] by Newton iteration on P_J in
// extended precision, w = 2/(J P_{J-1}(x) P'_J(x)), symmetrised and
// normalised (the construction numpy.polynomial.legendre.leggauss
// publishes), sorted north first and halved to sum to one
// (spectral.py:57-61).
void gauss_legendre(int J, std::vector<double> &mu, std::vector<double> &gw) {
  typedef long double R;
  std::vector<R> x(J), w(J);
  const R pi = 3.14159265358979323846264338327950288L;
  for (int i = 0; i < J; ++i) {
    R z = std::cos(pi * (i + 0.75L) / (J + 0.5L));
    R p0 = 1, p1 = z, dp = 0;
    for (int it = 0; it < 100; ++it) {
      p0 = 1;
      p1 = z;
      for (int n = 2; n <= J; ++n) {
        const R p2 = ((2 * n - 1) * z * p1 - (n - 1) * p0) / n;
        p0 = p1;
        p1 = p2;
      }
      dp = J * (z * p1 - p0) / (z * z - 1);
      const R dz = p1 / dp;
      z -= dz;
      if (std::fabs((double)dz) < 1e-19 && it > 2) break;
    }
    p0 = 1;
    p1 = z;
    for (int n = 2; n <= J; ++n) {
      const R p2 = ((2 * n - 1) * z * p1 - (n - 1) * p0) / n;
      p0 = p1;
      p1 = p2;
    }
    dp = J * (z * p1 - p0) / (z * z - 1);
    x[i] = z;  // descending: i = 0 is the northernmost root
    w[i] = 2 / (J * p0 * dp);
  }
  R sum = 0;
  std::vector<R> ws(J), xs(J);
  for (int i = 0; i < J; ++i) {
    ws[i] = (w[i] + w[J - 1 - i]) / 2;
    xs[i] = (x[i] - x[J - 1 - i]) / 2;
    sum += ws[i];
  }
  mu.assign(J, 0.0);
  gw.assign(J, 0.0);
  for (int i = 0; i < J; ++i) {
    mu[i] = (double)xs[i];
    gw[i] = (double)(ws[i] / sum);  // sum(gw) = 1
  }
}

// P[m][n][j] (n <= nmax+1) and H[m][n][j] (n <= nmax) of spectral.py:93-138.
void legendre(int nmax, const std::vector<double> &mu, std::vector<double> &P,
              std::vector<double> &H) {
  const int J = (int)mu.size(), M = nmax + 1, NX = nmax + 2;
  std::vector<double> eps((size_t)M * NX, 0.0), p((size_t)M * NX * J, 0.0);
  auto E = [&](int m, int n) -> double & { return eps[(size_t)m * NX + n]; };
  auto Pp = [&](int m, int n, int j) -> double & { return p[((size_t)m * NX + n) * J + j]; };
  for (int m = 0; m < M; ++m)
    for (int n = m + 1; n < NX; ++n)
      E(m, n) = std::sqrt((double(n) * n - double(m) * m) / (4.0 * n * n - 1.0));
  for (int j = 0; j < J; ++j) {
    const double x = mu[j], c = std::sqrt(1.0 - x * x);
    Pp(0, 0, j) = 1.0;
    for (int m = 1; m < M; ++m) Pp(m, m, j) = Pp(m - 1, m - 1, j) * std::sqrt((2.0 * m + 1.0) / (2.0 * m)) * c;
    for (int m = 0; m < M; ++m) {
      Pp(m, m + 1, j) = Pp(m, m, j) * std::sqrt(2.0 * m + 3.0) * x;
      for (int n = m + 2; n < NX; ++n)
        Pp(m, n, j) = (x * Pp(m, n - 1, j) - E(m, n - 1) * Pp(m, n - 2, j)) / E(m, n);
    }
  }
  P.assign((size_t)M * M * J, 0.0);
  H.assign((size_t)M * M * J, 0.0);
  for (int m = 0; m < M; ++m)
    for (int n = m; n < M; ++n)
      for (int j = 0; j < J; ++j) {
        P[((size_t)m * M + n) * J + j] = Pp(m, n, j);
        double h = -n * E(m, n + 1) * Pp(m, n + 1, j);
        if (n - 1 >= m) h += (n + 1.0) * E(m, n) * Pp(m, n - 1, j);
        H[((size_t)m * M + n) * J + j] = h;
      }
}

// In-place Gauss-Jordan inverse with partial pivoting (FP64).
bool invert(std::vector<double> &a, int n) {
  std::vector<double> inv((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[(size_t)r * n + c]) > std::fabs(a[(size_t)piv * n + c])) piv = r;
    if (a[(size_t)piv * n + c] == 0.0) return false;
    if (piv != c)
      for (int k = 0; k < n; ++k) {
        std::swap(a[(size_t)c * n + k], a[(size_t)piv * n + k]);
        std::swap(inv[(size_t)c * n + k], inv[(size_t)piv * n + k]);
      }
    const double d = 1.0 / a[(size_t)c * n + c];
    for (int k = 0; k < n; ++k) {
      a[(size_t)c * n + k] *= d;
      inv[(size_t)c * n + k] *= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = a[(size_t)r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        a[(size_t)r * n + k] -= f * a[(size_t)c * n + k];
        inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
      }
    }
  }
  a.swap(inv);
  return true;
}

template <class T>
struct DevTabs {
  T *Ps = nullptr, *Hs = nullptr, *Pa = nullptr, *Ha = nullptr;  // [s][Jh]
  T *inv_lap = nullptr, *nn1a2 = nullptr, *lap = nullptr;        // [M]
  T *rcos = nullptr;                                             // [J]
  T *vc = nullptr;                                               // VOff
  T *kv = nullptr, *kd = nullptr, *ktt = nullptr, *kq = nullptr; // [M][L]
  T *tref = nullptr;                                             // [nb][L]
  T *minv = nullptr;                                             // [nset][nb][M][L][L]
  T *vinv = nullptr;                                             // [nset][4][L][L]
  T *twa = nullptr;  // [I] complex exp(-2 pi i k / I): FFTs of the Stockham path
  // split-BF16x3 plans: Ps, Hs, Pa, Ha as BF16 high / low parts
  short *split_hi[4] = {}, *split_lo[4] = {};
};

// hi = bf16(x), lo = bf16(x - hi), both round to nearest even, of the FP32
// table value x (the rounding the kernels apply to their data operands)
inline short bf16_rne_host(float x) {
  unsigned u;
  std::memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
inline float bf16_up_host(short h) {
  const unsigned u = ((unsigned)(unsigned short)h) << 16;
  float x;
  std::memcpy(&x, &u, 4);
  return x;
}

template <class T>
int upload(T **dst, const std::vector<double> &src) {
  std::vector<T> tmp(src.begin(), src.end());
  if (*dst == nullptr) HIPCHK(hipMalloc((void **)dst, std::max<size_t>(tmp.size(), 1) * sizeof(T)));
  HIPCHK(hipMemcpy(*dst, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice));
  return LBM_OK;
}

template <class T>
void free_tabs(DevTabs<T> &t) {
  T *ps[] = {t.Ps, t.Hs, t.Pa, t.Ha, t.inv_lap, t.nn1a2, t.lap, t.rcos, t.vc,
             t.kv, t.kd, t.ktt, t.kq, t.tref, t.minv, t.vinv, t.twa};
  for (T *p : ps)
    if (p) (void)hipFree(p);
  for (int k = 0; k < 4; ++k) {
    if (t.split_hi[k]) (void)hipFree(t.split_hi[k]);
    if (t.split_lo[k]) (void)hipFree(t.split_lo[k]);
  }
  t = DevTabs<T>();
}

inline unsigned nblk(size_t n, unsigned bs = 256) {
  size_t b = (n + bs - 1) / bs;
  return (unsigned)std::min<size_t>(std::max<size_t>(b, 1), 65535u * 16u);
}

}  // namespace

// ---------------------------------------------------------------------------
struct lbm_plan {
  // sizes
  int J = 0, I = 0, N = 0, M = 0, L = 0, E = 0, nb = 0, prec = 0, nspec = 0, Jh = 0, nset = 0;
  int inistp = 2, strictly_linear = 0, skip_low_n = 1, vmask = 0, linear = 0;
  bool split = false;  // LBM_BF16X3: split-BF16x3 Legendre contractions
  // FFT lengths other than the specialised 128 and 384: the radices of the
  // Stockham passes (ffac.np = 0 for the specialised lengths)
  FftFactors ffac{};
  double a = 0, delt = 0, robert = 0;
  // host FP64 setup
  std::vector<double> mu, gw, coslat, lat, lon;
  std::vector<double> sigm, dsig, alpha, sig, dlog2, dlog3, cpg, w3, wP, dsigm, ww, lop;
  std::vector<double> kv, kd, ktt, kq;
  std::vector<double> vdif[4];
  std::vector<double> tbar, minv, vinv;
  // M(n) before inversion for the free time step (set inistp + 1), per basic state
  std::vector<double> mmat;
  double user_dt = 0;  // time step of matrix set inistp + 1 (lbm_prepare_dt)
  std::vector<short> n_of_s;
  // device
  DevTabs<double> d64;
  DevTabs<float> d32;
  short *d_n_of_s = nullptr;
  double *d_mu = nullptr;
  double *d_psbar = nullptr;  // [nb][J][I] FP64 basic surface pressure
  void *sa3 = nullptr, *sa2 = nullptr, *sb3 = nullptr, *tend3 = nullptr, *tend2 = nullptr;
  void *fin = nullptr, *gin = nullptr, *gout = nullptr, *fout = nullptr;
  void *raw = nullptr, *psa = nullptr, *psb = nullptr, *bsg = nullptr, *frc = nullptr;
  int nfrc = 0;
  void *sb2 = nullptr;  // [s][3][E] pi at t-dt in tangent-linear mode (field 0)
  // adjoint sweep: state adjoints of both time levels, synthesis adjoint of
  // the column sums, forcing adjoint
  void *aa3 = nullptr, *ab3 = nullptr, *aa2 = nullptr, *ab2 = nullptr, *s3a = nullptr;
  void *frc_a = nullptr;
  // linear response: vorticity functional, level and step weights, FP64 sums
  void *obs = nullptr, *wlev = nullptr, *wstep = nullptr;
  double *jacc = nullptr;
  int nwstep = 0, wstep_cap = 0;
  bool resp_set = false;
  // history: interval sums (two slots) and the last top sample
  int hist_mode = 0, hist_K = 0;
  bool keep_last = false;  // keep the last top sample (LBM.diagnostics)
  void *hspec[2] = {nullptr, nullptr}, *hgrid[2] = {nullptr, nullptr};
  void *hlast_spec = nullptr, *hlast_grid = nullptr;
  // MJO forcing: daily fields [f][day][j][L][I] interpolated into frc per step
  void *mjo = nullptr;
  int mjo_nd = 0, mjo_nt = 0, mjo_spin = 0, mjo_call0 = 0, mjo_single = 0;
  // steady vorticity/divergence [s][2][L][nsf] and ln ps [s][nsf] forcings (spectral)
  void *sfv = nullptr, *sfp = nullptr;
  int nsf = 0;
  // a run in progress (lbm_run_to): per-member lengths and steps done
  bool in_run = false;
  bool frc_in_run = false;  // the forcing changed between the steps of the last run
  std::vector<int> run_len;
  int run_done = 0;
  Ctl *ctl = nullptr;
  int *pend = nullptr, *d_nsteps = nullptr;
  std::vector<int> nsteps_h;
  bool basic_set = false, have_raw = false, raw_pending = false;
  // both time levels hold the same state (set by lbm_set_state or
  // lbm_set_spectral_state, until a step or a partial update separates them)
  bool levels_equal = true;
  long long steps_total = 0;
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  // a second copy of the prognostic state for the stand-alone tendency and
  // step entry points, which must leave the run's state untouched
  void *x_sa3 = nullptr, *x_sa2 = nullptr, *x_sb3 = nullptr, *x_sb2 = nullptr;
  void *x_psa = nullptr, *x_psb = nullptr, *x_raw = nullptr;
  // checkpointed trajectory of the default model for its adjoint: the
  // whole state every ckpt_every executed steps of a run (ckpt), the run it
  // belongs to (ckpt_len, ckpt_total; ckpt_ok once that run has ended), the
  // per-step trajectory of one recomputed segment (traj, ckpt_every slots)
  // and the slot the next tendency launch writes or the next transposed
  // launch reads (traj_cur, traj_step), the carried grid adjoints of ps'
  // (psa_a, psb_a) and the forward state kept aside during the sweep (keep)
  int ckpt_every = 0, ckpt_total = 0, ckpt_count = 0;
  bool ckpt_ok = false, ckpt_raw = false;
  std::vector<int> ckpt_len;
  void *ckpt = nullptr, *traj = nullptr, *keep = nullptr, *psa_a = nullptr, *psb_a = nullptr;
  size_t ckpt_cap = 0, traj_cap = 0;
  void *traj_cur = nullptr;
  int traj_step = -1;
  // raw grid (u, v, t, q) of the previous level after a partial update
  // (lbm_set_state_keep) that left the two levels different
  void *rawb = nullptr;
  int *x_pend = nullptr;
  Ctl *x_ctl = nullptr;

  hipStream_t own = nullptr;
  // second stream of the time step (launches beside the level kernels) and
  // its fork/join events: fft done, surface synthesis, MJO, history sums,
  // surface analysis, step start
  hipStream_t side = nullptr;
  hipEvent_t fj[6] = {};
  // graphs keyed by (kind, member count of the launches): kind 0 forward
  // chunk, 1 forward single step, 2 adjoint chunk, 3 adjoint single step
  std::map<std::pair<int, int>, hipGraphExec_t> graphs;
  // launch variants chosen by lbm_autotune (0 = the built-in choice):
  // inverse-FFT rows per block, forward-FFT rows per block, occupancy target
  // of the level synthesis and of the level tendency analyses
  static constexpr int NTUNE = 5;
  int tune[NTUNE] = {0, 0, 0, 0, 0};
  // kernel timing (lbm_profile)
  static constexpr int NPROF = 8;
  bool prof = false;
  hipEvent_t ev[NPROF + 1] = {};
  double prof_ms[NPROF] = {};
  long long prof_n[NPROF] = {};

  template <class T>
  DevTabs<T> &tabs();
  size_t rsize() const { return prec == LBM_FP64 ? 8 : 4; }
  hipStream_t strm(void *s) const { return s ? (hipStream_t)s : own; }
};
template <>
DevTabs<double> &lbm_plan::tabs<double>() { return d64; }
template <>
DevTabs<float> &lbm_plan::tabs<float>() { return d32; }

namespace {

int scratch_get(lbm_plan *p, size_t bytes, void **out, hipStream_t st) {
  if (bytes > p->scratch_bytes) {
    if (p->scratch) {
      HIPCHK(hipStreamSynchronize(st));
      HIPCHK(hipFree(p->scratch));
      p->scratch = nullptr;
    }
    HIPCHK(hipMalloc(&p->scratch, bytes));
    p->scratch_bytes = bytes;
  }
  *out = p->scratch;
  return LBM_OK;
}

void destroy_graphs(lbm_plan *p) {
  for (auto &kv : p->graphs) (void)hipGraphExecDestroy(kv.second);
  p->graphs.clear();
}

// Members that can still be live from step st of a run on (advancing, or,
// forward, with the pending ps' update of their last step), as a count of
// leading members rounded up to 8; E when E is not a multiple of 8.
int live_members(const lbm_plan *p, const std::vector<int> &len, int st, bool pending) {
  if (p->E % 8) return p->E;
  int last = -1;
  for (int e = 0; e < p->E; ++e)
    if (st < p->inistp + len[e] + (pending ? 1 : 0)) last = e;
  return std::max(8, (last + 1 + 7) / 8 * 8);
}

// every member advances through steps lo..hi (no mask needed there)
bool all_live(const lbm_plan *p, const std::vector<int> &len, int hi) {
  for (int e = 0; e < p->E; ++e)
    if (hi >= p->inistp + len[e]) return false;
  return true;
}

// launch extents of the column-tiled kernels for a mask
unsigned tiles_x(int ncols, const ActiveMask &am) {
  return am.reduced() ? (unsigned)((ncols / am.E) * (am.ecnt / 8)) : (unsigned)((ncols + 7) / 8);
}
long fft_rows(long J, int C, const ActiveMask &am) {
  return am.reduced() ? J * (C / am.E) * am.ecnt : J * fpad(C);
}

// Longitudes per block of the grid-point kernels (one thread each): 128, or
// 64 where the row length is not a multiple of 128.
int grid_chunk(int I) { return I % 128 == 0 ? 128 : 64; }

// Tables of one precision from the FP64 host setup.
template <class T>
int build_tabs(lbm_plan *p) {
  DevTabs<T> &t = p->tabs<T>();
  const int M = p->M, J = p->J, Jh = p->Jh, L = p->L;
  std::vector<double> P, H;
  legendre(p->N, p->mu, P, H);
  std::vector<double> Ps((size_t)p->nspec * Jh), Hs(Ps.size()), Pa(Ps.size()), Ha(Ps.size());
  for (int m = 0; m < M; ++m)
    for (int n = m; n < M; ++n) {
      const int s = spec_off(m, M) + n - m;
      for (int j = 0; j < Jh; ++j) {
        const double pv = P[((size_t)m * M + n) * J + j], hv = H[((size_t)m * M + n) * J + j];
        Ps[(size_t)s * Jh + j] = pv;
        Hs[(size_t)s * Jh + j] = hv;
        Pa[(size_t)s * Jh + j] = pv * p->gw[j];
        Ha[(size_t)s * Jh + j] = hv * p->gw[j];
      }
    }
  CHK(upload(&t.Ps, Ps));
  CHK(upload(&t.Hs, Hs));
  CHK(upload(&t.Pa, Pa));
  CHK(upload(&t.Ha, Ha));
  if (p->ffac.np > 0) {
    std::vector<double> tw(2 * (size_t)p->I);
    for (int k = 0; k < p->I; ++k) {
      const long double ang = -2.0L * (long double)M_PI * k / p->I;
      tw[2 * k] = (double)std::cos(ang);
      tw[2 * k + 1] = (double)std::sin(ang);
    }
    CHK(upload(&t.twa, tw));
  }
  if (std::is_same<T, float>::value && p->split) {
    const std::vector<double> *tabs4[4] = {&Ps, &Hs, &Pa, &Ha};
    for (int k = 0; k < 4; ++k) {
      std::vector<short> hi(tabs4[k]->size()), lo(hi.size());
      for (size_t i = 0; i < hi.size(); ++i) {
        const float x = (float)(*tabs4[k])[i];
        hi[i] = bf16_rne_host(x);
        lo[i] = bf16_rne_host(x - bf16_up_host(hi[i]));
      }
      HIPCHK(hipMalloc((void **)&t.split_hi[k], hi.size() * sizeof(short)));
      HIPCHK(hipMalloc((void **)&t.split_lo[k], lo.size() * sizeof(short)));
      HIPCHK(hipMemcpy(t.split_hi[k], hi.data(), hi.size() * sizeof(short), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(t.split_lo[k], lo.data(), lo.size() * sizeof(short), hipMemcpyHostToDevice));
    }
  }
  std::vector<double> il(M), nn(M), lp(M), rc(J);
  for (int n = 0; n < M; ++n) {
    const double nn1 = double(n) * (n + 1.0);
    il[n] = n > 0 ? -p->a * p->a / nn1 : 0.0;
    nn[n] = nn1 / (p->a * p->a);
    lp[n] = -nn1 / (p->a * p->a);
  }
  for (int j = 0; j < J; ++j) rc[j] = 1.0 / (p->a * p->coslat[j]);
  CHK(upload(&t.inv_lap, il));
  CHK(upload(&t.nn1a2, nn));
  CHK(upload(&t.lap, lp));
  CHK(upload(&t.rcos, rc));
  VOff vo{L};
  std::vector<double> vc(vo.size());
  for (int k = 0; k < L; ++k) {
    vc[vo.dsig() + k] = p->dsig[k];
    vc[vo.alpha() + k] = p->alpha[k];
    vc[vo.w3() + k] = p->w3[k];
    vc[vo.wP() + k] = p->wP[k];
    vc[vo.cpg() + k] = p->cpg[k];
    vc[vo.dlog3() + k] = p->dlog3[k];
    vc[vo.sig() + k] = p->sig[k];
  }
  for (int k = 0; k <= L; ++k) vc[vo.sigm() + k] = p->sigm[k];
  CHK(upload(&t.vc, vc));
  CHK(upload(&t.kv, p->kv));
  CHK(upload(&t.kd, p->kd));
  CHK(upload(&t.ktt, p->ktt));
  CHK(upload(&t.kq, p->kq));
  return LBM_OK;
}

// exp(-2 pi i k / n), k < n, for the 128- and 384-point transforms
template <class T>
int set_twiddles() {
  for (int n : {128, 384}) {
    std::vector<cx<T>> tw(n);
    for (int k = 0; k < n; ++k) {
      const long double ang = -2.0L * (long double)M_PI * k / n;
      tw[k] = mk((T)std::cos(ang), (T)std::sin(ang));
    }
    const size_t b = sizeof(cx<T>) * n;
    if (sizeof(T) == 8)
      HIPCHK(n == 128 ? hipMemcpyToSymbol(HIP_SYMBOL(c_tw128_d), tw.data(), b)
                      : hipMemcpyToSymbol(HIP_SYMBOL(c_tw384_d), tw.data(), b));
    else
      HIPCHK(n == 128 ? hipMemcpyToSymbol(HIP_SYMBOL(c_tw128_f), tw.data(), b)
                      : hipMemcpyToSymbol(HIP_SYMBOL(c_tw384_f), tw.data(), b));
  }
  return LBM_OK;
}

// ---------------------------------------------------------------------------
// Transform building blocks on internal buffers.
// ---------------------------------------------------------------------------
// Synthesis with the synthesis tables P, H and -a^2/(n(n+1)) by default; the
// transposed analysis passes the analysis tables instead.
// The pre-split halves of a plan table (Ps, Hs, Pa or Ha) pair.
template <class T>
SplitTabs split_of(DevTabs<T> &t, const T *P, const T *H) {
  SplitTabs sp;
  const T *tabs4[4] = {t.Ps, t.Hs, t.Pa, t.Ha};
  for (int k = 0; k < 4; ++k) {
    if (tabs4[k] == P) sp.Ph = t.split_hi[k], sp.Pl = t.split_lo[k];
    if (tabs4[k] == H) sp.Hh = t.split_hi[k], sp.Hl = t.split_lo[k];
  }
  return sp;
}

// Recipes whose fields stay in plain FP32 in split-BF16x3 plans: the
// surface fields (pi, its gradient and the column sums) and every scalar
// analysis.  pi is analysed again from the ps' grid at every step, so its
// transform error accumulates over the run instead of staying a per-step
// rounding; these contractions are a few columns per member.
template <class R>
constexpr bool split_allowed() {
  return !std::is_same<R, SynSurf>::value && !std::is_same<R, AnaScalar>::value &&
         !std::is_same<R, AdjSynSurf>::value && !std::is_same<R, AdjScalarFft>::value;
}

// Occupancy targets (waves per SIMD) of the autotuned Legendre kernels:
// variant 0 is the built-in one.
constexpr int SYN_OCC[3] = {3, 2, 4};
constexpr int ANA_OCC[3] = {4, 3, 2};

template <class T, class R, int OCC>
int launch_syn_k(lbm_plan *p, const void *in, void *out, int ncols, int C, OutSlots sl,
                 hipStream_t st, int in_nf, const T *P, const T *H, const T *per_n, ActiveMask am) {
  DevTabs<T> &t = p->tabs<T>();
  dim3 grid(tiles_x(ncols, am), p->M, (p->Jh + 31) / 32);
  const T *tP = P ? P : t.Ps, *tH = H ? H : t.Hs;
  if constexpr (std::is_same<T, float>::value && split_allowed<R>()) {
    if (p->split) {
      hipLaunchKernelGGL((k_syn_mfma<T, R, true, OCC>), grid, dim3(256), 0, st, (const cx<T> *)in,
                         (cx<T> *)out, tP, tH, per_n ? per_n : t.inv_lap, t.rcos, ncols, C, sl,
                         p->J, p->M, in_nf, am, split_of<T>(t, tP, tH));
      HIPCHK(hipGetLastError());
      return LBM_OK;
    }
  }
  hipLaunchKernelGGL((k_syn_mfma<T, R, false, OCC>), grid, dim3(256), 0, st, (const cx<T> *)in,
                     (cx<T> *)out, tP, tH, per_n ? per_n : t.inv_lap, t.rcos, ncols, C, sl, p->J,
                     p->M, in_nf, am, SplitTabs{});
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

// Synthesis with the synthesis tables P, H and -a^2/(n(n+1)) by default; the
// transposed analysis passes the analysis tables instead.  The level
// synthesis of the time step runs the autotuned occupancy variant.
template <class T, class R>
int launch_syn(lbm_plan *p, const void *in, void *out, int ncols, int C, OutSlots sl,
               hipStream_t st, int in_nf = R::NIN, const T *P = nullptr, const T *H = nullptr,
               const T *per_n = nullptr, ActiveMask am = ActiveMask{nullptr, nullptr, nullptr, 0, 1, 0}) {
  if constexpr (std::is_same<R, SynLevel>::value) {
    if (p->tune[2] == 1)
      return launch_syn_k<T, R, SYN_OCC[1]>(p, in, out, ncols, C, sl, st, in_nf, P, H, per_n, am);
    if (p->tune[2] == 2)
      return launch_syn_k<T, R, SYN_OCC[2]>(p, in, out, ncols, C, sl, st, in_nf, P, H, per_n, am);
  }
  return launch_syn_k<T, R, SYN_OCC[0]>(p, in, out, ncols, C, sl, st, in_nf, P, H, per_n, am);
}

template <class T, class R, int OCC>
int launch_ana_k(lbm_plan *p, const void *in, void *out, int ncols, int C, InSlots sl,
                 hipStream_t st, int out_nf, int out_f0, const T *P, const T *H, const T *per_n,
                 ActiveMask am) {
  DevTabs<T> &t = p->tabs<T>();
  // each block covers 2 n-tiles of 16 per parity (64 degrees); more for large N
  dim3 grid(tiles_x(ncols, am), p->M, (p->M + 63) / 64);
  const T *tP = P ? P : t.Pa, *tH = H ? H : t.Ha;
  if constexpr (std::is_same<T, float>::value && split_allowed<R>()) {
    if (p->split) {
      hipLaunchKernelGGL((k_ana_mfma<T, R, true, OCC>), grid, dim3(256), 0, st, (const cx<T> *)in,
                         (cx<T> *)out, tP, tH, per_n ? per_n : t.nn1a2, ncols, C, sl, p->J, p->M,
                         out_nf, out_f0, am, split_of<T>(t, tP, tH));
      HIPCHK(hipGetLastError());
      return LBM_OK;
    }
  }
  hipLaunchKernelGGL((k_ana_mfma<T, R, false, OCC>), grid, dim3(256), 0, st, (const cx<T> *)in,
                     (cx<T> *)out, tP, tH, per_n ? per_n : t.nn1a2, ncols, C, sl, p->J, p->M,
                     out_nf, out_f0, am, SplitTabs{});
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

// Analysis with the analysis tables gw P, gw H and n(n+1)/a^2 by default;
// the transposed synthesis passes the synthesis tables instead.  The level
// tendency analyses of the time step run the autotuned occupancy variant.
template <class T, class R>
int launch_ana(lbm_plan *p, const void *in, void *out, int ncols, int C, InSlots sl,
               hipStream_t st, int out_nf = R::NOUT, int out_f0 = 0, const T *P = nullptr,
               const T *H = nullptr, const T *per_n = nullptr,
               ActiveMask am = ActiveMask{nullptr, nullptr, nullptr, 0, 1, 0}) {
  if constexpr (std::is_same<R, AnaTendVD>::value || std::is_same<R, AnaFluxDiv>::value) {
    if (p->tune[3] == 1)
      return launch_ana_k<T, R, ANA_OCC[1]>(p, in, out, ncols, C, sl, st, out_nf, out_f0, P, H,
                                            per_n, am);
    if (p->tune[3] == 2)
      return launch_ana_k<T, R, ANA_OCC[2]>(p, in, out, ncols, C, sl, st, out_nf, out_f0, P, H,
                                            per_n, am);
  }
  return launch_ana_k<T, R, ANA_OCC[0]>(p, in, out, ncols, C, sl, st, out_nf, out_f0, P, H, per_n,
                                        am);
}

// Rows per block of the autotuned FFT variants 1 and 2 (variant 0: the
// FftShape default): the LDS of a row bounds the blocks per CU, the rows
// per block the waves per barrier.
template <class T, int NI, bool INV>
constexpr int fft_alt_rows(int v) {
  if (NI == 128) return sizeof(T) == 8 ? (v == 1 ? 8 : 32) : (v == 1 ? 16 : 64);
  if (sizeof(T) == 4) return v == 1 ? 16 : 20;
  return INV ? (v == 1 ? 16 : 20) : (v == 1 ? 10 : 16);
}

template <class T, int NI, int RW>
int ifft_rows_launch(lbm_plan *p, const void *F, void *G, ColMap cm, hipStream_t st, int mode,
                     ActiveMask am) {
  using S = FftShape<T, NI, true, RW>;
  const int C = cm.ncols();
  const long rows = fft_rows(p->J, C, am);
  hipLaunchKernelGGL((k_ifft<T, NI, RW>), dim3((unsigned)((rows + S::ROWS - 1) / S::ROWS)),
                     dim3(S::THREADS), 0, st, (const cx<T> *)F, (T *)G, C, cm, p->J, p->M, mode, am);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

template <class T, int NI, int RW>
int fft_rows_launch(lbm_plan *p, const void *G, void *F, ColMap cm, hipStream_t st, int mode,
                    ActiveMask am) {
  using S = FftShape<T, NI, false, RW>;
  const int C = cm.ncols();
  const long rows = fft_rows(p->J, C, am);
  hipLaunchKernelGGL((k_fft<T, NI, RW>), dim3((unsigned)((rows + S::ROWS - 1) / S::ROWS)),
                     dim3(S::THREADS), 0, st, (const T *)G, (cx<T> *)F, C, cm, p->J, p->M, mode, am);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

template <class T, int NI>
int launch_ifft_n(lbm_plan *p, const void *F, void *G, ColMap cm, hipStream_t st, int mode,
                  ActiveMask am) {
  if (p->tune[0] == 1)
    return ifft_rows_launch<T, NI, fft_alt_rows<T, NI, true>(1)>(p, F, G, cm, st, mode, am);
  if (p->tune[0] == 2)
    return ifft_rows_launch<T, NI, fft_alt_rows<T, NI, true>(2)>(p, F, G, cm, st, mode, am);
  return ifft_rows_launch<T, NI, 0>(p, F, G, cm, st, mode, am);
}

template <class T, int NI>
int launch_fft_n(lbm_plan *p, const void *G, void *F, ColMap cm, hipStream_t st, int mode,
                 ActiveMask am) {
  if (p->tune[1] == 1)
    return fft_rows_launch<T, NI, fft_alt_rows<T, NI, false>(1)>(p, G, F, cm, st, mode, am);
  if (p->tune[1] == 2)
    return fft_rows_launch<T, NI, fft_alt_rows<T, NI, false>(2)>(p, G, F, cm, st, mode, am);
  return fft_rows_launch<T, NI, 0>(p, G, F, cm, st, mode, am);
}

// Rows per block of the Stockham FFTs (32 threads and 2 NC complex of LDS
// per row): one tile of 8 Fourier columns.
int fft_any_rows(int) { return FFT_ANY_ROWS; }

template <class T, bool INV>
int launch_fft_any(lbm_plan *p, const void *in, void *out, ColMap cm, hipStream_t st, int mode,
                   ActiveMask am) {
  const int C = cm.ncols(), rb = fft_any_rows(p->ffac.NC);
  const long rows = fft_rows(p->J, C, am);
  const dim3 grid((unsigned)((rows + rb - 1) / rb)), block(rb * FFT_ANY_TPR);
  const size_t lds = ((size_t)rb * 2 * p->ffac.NC + p->ffac.NI) * sizeof(cx<T>);
  const cx<T> *tw = (const cx<T> *)p->tabs<T>().twa;
  if (INV)
    hipLaunchKernelGGL((k_ifft_any<T>), grid, block, lds, st, (const cx<T> *)in, (T *)out, C, cm, p->J,
                       p->M, mode, am, p->ffac, tw);
  else
    hipLaunchKernelGGL((k_fft_any<T>), grid, block, lds, st, (const T *)in, (cx<T> *)out, C, cm, p->J,
                       p->M, mode, am, p->ffac, tw);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

template <class T>
int launch_ifft(lbm_plan *p, const void *F, void *G, ColMap cm, hipStream_t st, int mode = 0,
                ActiveMask am = ActiveMask{nullptr, nullptr, nullptr, 0, 1, 0}) {
  if (p->ffac.np > 0) return launch_fft_any<T, true>(p, F, G, cm, st, mode, am);
  return p->I == 128 ? launch_ifft_n<T, 128>(p, F, G, cm, st, mode, am)
                     : launch_ifft_n<T, 384>(p, F, G, cm, st, mode, am);
}

template <class T>
int launch_fft(lbm_plan *p, const void *G, void *F, ColMap cm, hipStream_t st, int mode = 0,
               ActiveMask am = ActiveMask{nullptr, nullptr, nullptr, 0, 1, 0}) {
  if (p->ffac.np > 0) return launch_fft_any<T, false>(p, G, F, cm, st, mode, am);
  return p->I == 128 ? launch_fft_n<T, 128>(p, G, F, cm, st, mode, am)
                     : launch_fft_n<T, 384>(p, G, F, cm, st, mode, am);
}

// members still advancing at the device step counter of the time loop
// (Ctl::step is the first field of the control block); the adjoint has no
// deferred ps' updates
static_assert(offsetof(Ctl, step) == 0, "the mask reads the step counter at the block's start");
// ecnt <= 0: every member is live through the launches (no mask at all)
ActiveMask loop_mask(const lbm_plan *p, bool with_pending, int ecnt) {
  if (ecnt <= 0) return ActiveMask{nullptr, nullptr, nullptr, 0, p->E, 0};
  return ActiveMask{reinterpret_cast<const int *>(p->ctl), p->d_nsteps,
                    with_pending ? p->pend : nullptr, p->inistp, p->E,
                    ecnt < p->E ? ecnt : 0};
}

// ---------------------------------------------------------------------------
// Spectral API: (B, nk) transforms on caller FP64 buffers, and their
// transposes (the adjoint twins, for the Euclidean inner products of the
// grid values and of the re/im parts of the n >= m coefficients).
// ---------------------------------------------------------------------------
enum SpecOp {
  OP_G2W, OP_W2G, OP_UV2VD, OP_PC2UV, OP_HGRAD,
  OP_G2W_ADJ, OP_W2G_ADJ, OP_UV2VD_ADJ, OP_PC2UV_ADJ, OP_HGRAD_ADJ
};

template <class T>
int spectral_op(lbm_plan *p, SpecOp op, const double *in0, const double *in1, double *out0,
                double *out1, int B, int nk, hipStream_t st) {
  if (B <= 0 || nk <= 0) return fail(LBM_ERR_VALUE, "batch and column counts must be positive");
  const int J = p->J, I = p->I, M = p->M;
  const int ncols = B * nk;
  // two fields of Fourier columns per wavenumber, 32-bit offsets in the analysis
  if ((size_t)J * ((2 * (size_t)ncols + 7) & ~(size_t)7) * sizeof(cx<T>) >= ((size_t)1 << 32))
    return fail(LBM_ERR_VALUE, "batch %d x %d columns: one wavenumber of the Fourier columns exceeds 4 GiB",
                B, nk);
  // grid -> spectral (analysis kernels) or spectral -> grid (synthesis kernels)
  const bool to_spec = op == OP_G2W || op == OP_UV2VD || op == OP_W2G_ADJ || op == OP_PC2UV_ADJ ||
                       op == OP_HGRAD_ADJ;
  const bool adj = op >= OP_G2W_ADJ;
  int ngrid, nspecf;  // grid and spectral fields of the call
  switch (op) {
    case OP_G2W: case OP_W2G: case OP_G2W_ADJ: case OP_W2G_ADJ: ngrid = 1; nspecf = 1; break;
    case OP_HGRAD: case OP_HGRAD_ADJ: ngrid = 2; nspecf = 1; break;
    default: ngrid = 2; nspecf = 2; break;
  }
  const size_t gbytes = (size_t)B * J * ngrid * nk * I * sizeof(T);
  const size_t fbytes = (size_t)M * J * fpad(ngrid * ncols) * sizeof(cx<T>);
  const size_t sbytes = (size_t)p->nspec * nspecf * ncols * sizeof(cx<T>);
  const size_t al = 256;
  auto up = [&](size_t b) { return (b + al - 1) / al * al; };
  void *base;
  CHK(scratch_get(p, up(gbytes) + up(fbytes) + up(sbytes), &base, st));
  char *cb = (char *)base;
  T *g = (T *)cb;
  cx<T> *f = (cx<T> *)(cb + up(gbytes));
  cx<T> *s = (cx<T> *)(cb + up(gbytes) + up(fbytes));
  ColMap cm{B, nk, ngrid, 0};
  const size_t gtot = (size_t)B * J * I * nk;
  const double *rc64 = p->d64.rcos;
  const size_t stot = (size_t)p->nspec * ncols;
  const size_t spec_out = (size_t)B * M * M * nk;
  DevTabs<T> &t = p->tabs<T>();
  if (to_spec) {
    // vector grid inputs carry 1/(a cos lat): the forward's pre-scale, or
    // the transpose of a scaled synthesis output
    const bool scale = op != OP_G2W && op != OP_W2G_ADJ;
    const double *gin[2] = {in0, in1};
    for (int k = 0; k < ngrid; ++k)
      hipLaunchKernelGGL((k_pack_grid<T>), dim3(nblk(gtot)), dim3(256), 0, st, gin[k], g, B, J, I,
                         nk, ngrid * nk, k * nk, scale ? rc64 : (const double *)nullptr);
    HIPCHK(hipGetLastError());
    CHK(launch_fft<T>(p, g, f, cm, st, adj ? 2 : 0));
    InSlots sl{{0, ncols}};
    const int C = ngrid * ncols;
    switch (op) {
      case OP_G2W: CHK((launch_ana<T, AnaScalar>(p, f, s, ncols, C, sl, st))); break;
      case OP_UV2VD: CHK((launch_ana<T, AnaVorDiv>(p, f, s, ncols, C, sl, st))); break;
      case OP_W2G_ADJ:
        CHK((launch_ana<T, AnaScalar>(p, f, s, ncols, C, sl, st, 1, 0, t.Ps, t.Hs, t.inv_lap)));
        break;
      case OP_PC2UV_ADJ:
        CHK((launch_ana<T, AdjPsiChi>(p, f, s, ncols, C, sl, st, 2, 0, t.Ps, t.Hs, t.inv_lap)));
        break;
      default:
        CHK((launch_ana<T, AdjGrad>(p, f, s, ncols, C, sl, st, 1, 0, t.Ps, t.Hs, t.inv_lap)));
        break;
    }
    double *sout[2] = {out0, out1};
    for (int k = 0; k < nspecf; ++k)
      hipLaunchKernelGGL((k_unpack_spec<T>), dim3(nblk(spec_out)), dim3(256), 0, st, s, sout[k], B,
                         M, nk, nspecf, k, ncols);
    HIPCHK(hipGetLastError());
    return LBM_OK;
  }
  // synthesis family
  const double *sin_[2] = {in0, in1};
  for (int k = 0; k < nspecf; ++k)
    hipLaunchKernelGGL((k_pack_spec<T>), dim3(nblk(stot)), dim3(256), 0, st, sin_[k], s, B, M, nk,
                       nspecf, k, ncols);
  HIPCHK(hipGetLastError());
  OutSlots sl{{0, ncols}};
  const int C = ngrid * ncols;
  switch (op) {
    case OP_W2G: CHK((launch_syn<T, SynScalar>(p, s, f, ncols, C, sl, st))); break;
    case OP_PC2UV: CHK((launch_syn<T, SynPsiChi>(p, s, f, ncols, C, sl, st))); break;
    case OP_HGRAD: CHK((launch_syn<T, SynGrad>(p, s, f, ncols, C, sl, st))); break;
    case OP_G2W_ADJ:
      CHK((launch_syn<T, SynScalar>(p, s, f, ncols, C, sl, st, 1, t.Pa, t.Ha, t.nn1a2)));
      break;
    default:
      CHK((launch_syn<T, AdjVorDiv>(p, s, f, ncols, C, sl, st, 2, t.Pa, t.Ha, t.nn1a2)));
      break;
  }
  CHK(launch_ifft<T>(p, f, g, cm, st, adj ? 1 : 0));
  double *gout[2] = {out0, out1};
  for (int k = 0; k < ngrid; ++k)
    hipLaunchKernelGGL((k_unpack_grid<T>), dim3(nblk(gtot)), dim3(256), 0, st, g, gout[k], B, J, I,
                       nk, ngrid * nk, k * nk, 0, (const T *)nullptr, 0, 0, 1);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

int spectral_dispatch(lbm_plan *p, SpecOp op, const double *a0, const double *a1, double *o0,
                      double *o1, int B, int nk, void *stream) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  hipStream_t st = p->strm(stream);
  if (p->prec == LBM_FP64) return spectral_op<double>(p, op, a0, a1, o0, o1, B, nk, st);
  return spectral_op<float>(p, op, a0, a1, o0, o1, B, nk, st);
}

// ---------------------------------------------------------------------------
// Vertical operators (vertical.py:59-129) and the vertical-diffusion
// Laplacian (integrator.py:182-206), FP64 on the host.
// ---------------------------------------------------------------------------
int build_vertical(lbm_plan *p, const double *sigm) {
  const int L = p->L;
  p->sigm.assign(sigm, sigm + L + 1);
  p->dsig.resize(L);
  p->alpha.resize(L);
  p->sig.resize(L);
  p->dlog2.resize(L);
  p->dlog3.resize(L);
  p->cpg.resize(L);
  p->w3.resize(L);
  p->wP.resize(L);
  for (int k = 0; k < L; ++k) {
    const double bot = sigm[k], top = sigm[k + 1];
    if (!(bot > top) || top < 0.0) return fail(LBM_ERR_VALUE, "sigm must decrease from 1 to 0");
    p->dsig[k] = bot - top;
    const double ratio = top > 0.0 ? top * (std::log(bot) - std::log(top)) / (bot - top) : 0.0;
    p->alpha[k] = 1.0 - ratio;
    p->sig[k] = bot * std::exp(-p->alpha[k]);
    p->dlog2[k] = top > 0.0 ? std::log(p->sig[k]) - std::log(top) : 0.0;
    p->dlog3[k] = top > 0.0 ? std::log(bot) - std::log(top) : 0.0;
    p->cpg[k] = (bot * p->alpha[k] + top * p->dlog2[k]) / p->dsig[k];
    p->w3[k] = p->dlog3[k] / p->dsig[k];
    p->wP[k] = p->alpha[k] - p->cpg[k];
  }
  p->dsigm.assign(L + 1, 0.0);
  p->dsigm[0] = sigm[0] - p->sig[0];
  for (int k = 1; k < L; ++k) p->dsigm[k] = p->sig[k - 1] - p->sig[k];
  p->dsigm[L] = p->sig[L - 1] - sigm[L];
  p->ww.assign((size_t)L * L, 0.0);
  for (int k = 0; k < L; ++k) {
    for (int l = 0; l < k; ++l) p->ww[(size_t)k * L + l] = R_DRY * p->dlog3[l];
    p->ww[(size_t)k * L + k] = R_DRY * p->alpha[k];
  }
  // tridiagonal sigma Laplacian with zero-flux ends
  std::vector<double> lo(L, 0.0), up(L, 0.0), di(L, 0.0);
  const std::vector<double> &ds = p->dsig, &dm = p->dsigm;
  for (int k = 1; k < L; ++k) lo[k] = 1.0 / (ds[k] * dm[k]);
  for (int k = 0; k < L - 1; ++k) up[k] = 1.0 / (ds[k] * dm[k + 1]);
  for (int k = 0; k < L; ++k) di[k] = -(lo[k] + up[k]);
  up[0] = (dm[0] + dm[1]) / (ds[0] * dm[0] * dm[1]);
  lo[0] = 0.0;
  di[0] = -up[0];
  lo[L - 1] = (dm[L - 1] + dm[L]) / (ds[L - 1] * dm[L - 1] * dm[L]);
  up[L - 1] = 0.0;
  di[L - 1] = -lo[L - 1];
  p->lop.assign((size_t)L * L, 0.0);
  for (int k = 0; k < L; ++k) {
    p->lop[(size_t)k * L + k] = di[k];
    if (k > 0) p->lop[(size_t)k * L + k - 1] = lo[k];
    if (k < L - 1) p->lop[(size_t)k * L + k + 1] = up[k];
  }
  return LBM_OK;
}

// ---------------------------------------------------------------------------
// Basic state: FP64 transform pipeline, then the reference-temperature
// operators and the FP64 semi-implicit inverses on the host.
// ---------------------------------------------------------------------------
double step_dt(const lbm_plan *p, int set) {
  if (set > p->inistp) return p->user_dt;
  return set < p->inistp ? p->delt / std::ldexp(1.0, p->inistp - set) : p->delt;
}

// M(n) and its FP64 inverse for matrix sets lo..hi-1 (all by default)
int build_matrices(lbm_plan *p, int lo = 0, int hi = -1) {
  const int M = p->M, L = p->L, nb = p->nb, nset = p->nset;
  if (hi < 0) hi = nset;
  p->minv.resize((size_t)nset * nb * M * L * L, 0.0);
  p->mmat.resize((size_t)nb * M * L * L, 0.0);
  for (int set = lo; set < hi; ++set) {
    const double dt = step_dt(p, set), dt2 = 2.0 * dt;
    for (int b = 0; b < nb; ++b) {
      const double *tb = &p->tbar[(size_t)b * L];
      // tmtx (vertical.py:183-188), whmtx = WW tmtx, gcmtx = gg (x) dsig
      std::vector<double> tm((size_t)L * L, 0.0), wh((size_t)L * L, 0.0), gg(L);
      for (int k = 0; k < L; ++k) {
        tm[(size_t)k * L + k] = KAPPA * tb[k] * p->alpha[k];
        for (int l = k + 1; l < L; ++l) tm[(size_t)k * L + l] = KAPPA * tb[k] * p->w3[k] * p->dsig[l];
        gg[k] = R_DRY * tb[k] * p->cpg[k];
      }
      for (int r = 0; r < L; ++r)
        for (int c = 0; c < L; ++c) {
          double acc = 0;
          for (int q = 0; q < L; ++q) acc += p->ww[(size_t)r * L + q] * tm[(size_t)q * L + c];
          wh[(size_t)r * L + c] = acc;
        }
      // M(n) = diag(fd ft) - dt^2 lap_n (whmtx + gcmtx ft)  (integrator.py:149-158)
      for (int n = 0; n < M; ++n) {
        const double lap = -double(n) * (n + 1.0) / (p->a * p->a);
        std::vector<double> mat((size_t)L * L);
        for (int r = 0; r < L; ++r) {
          const double ft = 1.0 + dt2 * p->ktt[(size_t)n * L + r];
          const double fd = 1.0 + dt2 * p->kd[(size_t)n * L + r];
          for (int c = 0; c < L; ++c)
            mat[(size_t)r * L + c] = -dt * dt * lap * (wh[(size_t)r * L + c] + gg[r] * p->dsig[c] * ft);
          mat[(size_t)r * L + r] += fd * ft;
        }
        if (set == nset - 1)
          std::copy(mat.begin(), mat.end(), p->mmat.begin() + ((size_t)b * M + n) * L * L);
        if (!invert(mat, L))
          return fail(LBM_ERR_VALUE, "singular semi-implicit matrix (n=%d, basic state %d)", n, b);
        std::copy(mat.begin(), mat.end(), p->minv.begin() + (((size_t)set * nb + b) * M + n) * L * L);
      }
    }
  }
  return LBM_OK;
}

// (I - dt nu lop)^-1 per variable and dt (integrator.py:216-228)
int build_vdiff(lbm_plan *p, int lo = 0, int hi = -1) {
  const int L = p->L, nset = p->nset;
  if (hi < 0) hi = nset;
  p->vinv.resize((size_t)nset * 4 * L * L, 0.0);
  p->vmask = 0;
  for (int f = 0; f < 4; ++f)
    if (!p->vdif[f].empty()) p->vmask |= 1 << f;
  for (int set = lo; set < hi; ++set) {
    const double dt = step_dt(p, set);
    for (int f = 0; f < 4; ++f) {
      std::vector<double> mat((size_t)L * L, 0.0);
      for (int r = 0; r < L; ++r) {
        const double nu = p->vdif[f].empty() ? 0.0 : p->vdif[f][r];
        for (int c = 0; c < L; ++c)
          mat[(size_t)r * L + c] = (r == c ? 1.0 : 0.0) - dt * nu * p->lop[(size_t)r * L + c];
      }
      if (!invert(mat, L)) return fail(LBM_ERR_VALUE, "singular vertical-diffusion matrix");
      std::copy(mat.begin(), mat.end(), p->vinv.begin() + ((size_t)set * 4 + f) * L * L);
    }
  }
  return LBM_OK;
}

template <class T>
int upload_operators(lbm_plan *p) {
  DevTabs<T> &t = p->tabs<T>();
  CHK(upload(&t.tref, p->tbar));
  CHK(upload(&t.minv, p->minv));
  CHK(upload(&t.vinv, p->vinv));
  return LBM_OK;
}

int set_basic(lbm_plan *p, const double *u, const double *v, const double *t, const double *q,
              const double *ps, hipStream_t st) {
  const int J = p->J, I = p->I, L = p->L, nb = p->nb, M = p->M;
  const BRow br{L};
  const int ncl = nb * L;
  const ColMap cm{nb, L, 2, 1};  // u/(a cos), v/(a cos), ln ps
  const int C = cm.ncols();
  const size_t g1 = (size_t)nb * J * cm.rows() * I;
  const size_t f1 = (size_t)M * J * fpad(C);        // complex
  const size_t s1 = (size_t)p->nspec * C;           // complex
  const size_t g2 = (size_t)nb * J * 2 * L * I, g3 = (size_t)nb * J * 2 * I;
  const size_t bs = (size_t)nb * J * br.rows() * I;
  double *w = nullptr;
  const size_t total = g1 + 2 * f1 + 2 * s1 + g2 + g3 + bs + (size_t)nb * L;
  HIPCHK(hipMalloc(&w, total * sizeof(double)));
  double *cur = w;
  double *gtmp = cur; cur += g1;
  cx<double> *F = (cx<double> *)cur; cur += 2 * f1;
  cx<double> *S = (cx<double> *)cur; cur += 2 * s1;
  double *gs = cur; cur += g2;
  double *gg = cur; cur += g3;
  double *bsg64 = cur; cur += bs;
  double *tb = cur;
  const double *rc = p->d64.rcos;
  const size_t n3 = (size_t)nb * J * I * L, n2 = (size_t)nb * J * I;
  hipLaunchKernelGGL((k_pack_grid<double>), dim3(nblk(n3)), dim3(256), 0, st, u, gtmp, nb, J, I, L,
                     cm.rows(), 0, rc);
  hipLaunchKernelGGL((k_pack_grid<double>), dim3(nblk(n3)), dim3(256), 0, st, v, gtmp, nb, J, I, L,
                     cm.rows(), L, rc);
  hipLaunchKernelGGL((k_pack_grid<double>), dim3(nblk(n2)), dim3(256), 0, st, ps, gtmp, nb, J, I, 1,
                     cm.rows(), 2 * L, (const double *)nullptr);
  hipLaunchKernelGGL((k_log_row<double>), dim3(nblk(n2)), dim3(256), 0, st, gtmp, nb, J, I,
                     cm.rows(), 2 * L);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(p->d_psbar, ps, n2 * sizeof(double), hipMemcpyDeviceToDevice, st));
  CHK(launch_fft<double>(p, gtmp, F, cm, st));
  cx<double> *Sln = S + (size_t)p->nspec * 2 * ncl;
  {
    InSlots sl{{0, ncl}};
    CHK((launch_ana<double, AnaVorDiv>(p, F, S, ncl, C, sl, st)));
    InSlots s2{{2 * ncl}};
    CHK((launch_ana<double, AnaScalar>(p, F, Sln, nb, C, s2, st, 1, 0)));
  }
  // grid vor, div (two fields as 2*ncl columns of one) and grad ln ps
  {
    OutSlots sl{{0}};
    CHK((launch_syn<double, SynScalar>(p, S, F, 2 * ncl, 2 * ncl, sl, st)));
    CHK(launch_ifft<double>(p, F, gs, ColMap{nb, L, 2, 0}, st));
    OutSlots s2{{0, nb}};
    CHK((launch_syn<double, SynGrad>(p, Sln, F, nb, 2 * nb, s2, st)));
    CHK(launch_ifft<double>(p, F, gg, ColMap{nb, 1, 0, 2}, st));
  }
  hipLaunchKernelGGL(k_basic_terms, dim3(nb * J * (I / grid_chunk(I))), dim3(grid_chunk(I)), 0, st, u, v, t, q, ps, gs, gg,
                     p->d64.vc, p->d_mu, bsg64, J, I, L);
  hipLaunchKernelGGL(k_tbar, dim3(nb * L), dim3(256), 0, st, bsg64, tb, J, I, L);
  HIPCHK(hipGetLastError());
  p->tbar.assign((size_t)nb * L, 0.0);
  HIPCHK(hipMemcpyAsync(p->tbar.data(), tb, (size_t)nb * L * sizeof(double), hipMemcpyDeviceToHost, st));
  if (p->prec == LBM_FP64) {
    HIPCHK(hipMemcpyAsync(p->bsg, bsg64, bs * sizeof(double), hipMemcpyDeviceToDevice, st));
  } else {
    hipLaunchKernelGGL((k_convert<double, float>), dim3(nblk(bs)), dim3(256), 0, st, bsg64,
                       (float *)p->bsg, bs);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipFree(w));
  CHK(build_matrices(p));
  CHK(upload_operators<double>(p));
  if (p->prec != LBM_FP64) CHK(upload_operators<float>(p));
  p->basic_set = true;
  return LBM_OK;
}

// ---------------------------------------------------------------------------
// State I/O
// ---------------------------------------------------------------------------
template <class T>
int set_state_t(lbm_plan *p, const double *u, const double *v, const double *t, const double *q,
                const double *ps, hipStream_t st) {
  const int J = p->J, I = p->I, L = p->L, E = p->E, M = p->M;
  const int EL = E * L;
  const size_t n3 = (size_t)E * J * I * L, n2 = (size_t)E * J * I;
  const size_t s3 = (size_t)p->nspec * 4 * EL, s2 = (size_t)p->nspec * 3 * E;
  HIPCHK(hipMemsetAsync(p->pend, 0, E * sizeof(int), st));
  if (ps) {
    hipLaunchKernelGGL((k_pack_grid<T>), dim3(nblk(n2)), dim3(256), 0, st, ps, (T *)p->psa, E, J, I,
                       1, 1, 0, (const double *)nullptr);
    HIPCHK(hipGetLastError());
  } else {
    HIPCHK(hipMemsetAsync(p->psa, 0, n2 * sizeof(T), st));
  }
  HIPCHK(hipMemcpyAsync(p->psb, p->psa, n2 * sizeof(T), hipMemcpyDeviceToDevice, st));
  const double *src[4] = {u, v, t, q};
  if (!(u || v || t || q || ps)) {
    HIPCHK(hipMemsetAsync(p->sa3, 0, s3 * sizeof(cx<T>), st));
    HIPCHK(hipMemsetAsync(p->sb3, 0, s3 * sizeof(cx<T>), st));
    HIPCHK(hipMemsetAsync(p->sa2, 0, s2 * sizeof(cx<T>), st));
    if (p->sb2) HIPCHK(hipMemsetAsync(p->sb2, 0, s2 * sizeof(cx<T>), st));
    p->have_raw = false;
    p->raw_pending = false;
    return LBM_OK;
  }
  // raw grid (u, v, t, q) for the first tendency call
  for (int f = 0; f < 4; ++f) {
    if (src[f])
      hipLaunchKernelGGL((k_pack_grid<T>), dim3(nblk(n3)), dim3(256), 0, st, src[f], (T *)p->raw, E,
                         J, I, L, 4 * L, f * L, (const double *)nullptr);
    else
      hipLaunchKernelGGL((k_zero_rows<T>), dim3(nblk(n3)), dim3(256), 0, st, (T *)p->raw, E, J, I,
                         4 * L, f * L, L);
  }
  HIPCHK(hipGetLastError());
  // analysis (state.py:86-92): rows u/(a cos), v/(a cos), t, q; surface pi'
  const ColMap cm{E, L, 4, 1};
  const int C = cm.ncols();
  const size_t gb = ((size_t)E * J * cm.rows() * I * sizeof(T) + 255) / 256 * 256;
  const size_t fb = (size_t)M * J * fpad(C) * sizeof(cx<T>);
  void *base;
  CHK(scratch_get(p, gb + fb, &base, st));
  T *g = (T *)base;
  cx<T> *F = (cx<T> *)((char *)base + gb);
  const double *rc = p->d64.rcos;
  for (int f = 0; f < 4; ++f) {
    if (src[f])
      hipLaunchKernelGGL((k_pack_grid<T>), dim3(nblk(n3)), dim3(256), 0, st, src[f], g, E, J, I, L,
                         cm.rows(), f * L, f < 2 ? rc : (const double *)nullptr);
    else
      hipLaunchKernelGGL((k_zero_rows<T>), dim3(nblk(n3)), dim3(256), 0, st, g, E, J, I, cm.rows(),
                         f * L, L);
  }
  hipLaunchKernelGGL((k_pack_pi<T>), dim3(nblk(n2)), dim3(256), 0, st, ps, g, E, J, I, cm.rows(),
                     4 * L, (const double *)p->d_psbar, p->nb, p->linear);
  HIPCHK(hipGetLastError());
  CHK(launch_fft<T>(p, g, F, cm, st));
  {
    InSlots sl{{0, EL, 2 * EL, 3 * EL}};
    CHK((launch_ana<T, AnaState>(p, F, p->sa3, EL, C, sl, st)));
    InSlots s2{{4 * EL}};
    CHK((launch_ana<T, AnaScalar>(p, F, p->sa2, E, C, s2, st, 3, 0)));
  }
  hipLaunchKernelGGL((k_vsum<T>), dim3(nblk((size_t)p->nspec * E)), dim3(256), 0, st,
                     (const cx<T> *)p->sa3, (cx<T> *)p->sa2, p->tabs<T>().vc, E, L, p->nspec);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(p->sb3, p->sa3, s3 * sizeof(cx<T>), hipMemcpyDeviceToDevice, st));
  if (p->sb2) HIPCHK(hipMemcpyAsync(p->sb2, p->sa2, s2 * sizeof(cx<T>), hipMemcpyDeviceToDevice, st));
  p->have_raw = true;
  p->raw_pending = true;
  return LBM_OK;
}

// Synthesis of the current spectral level into gin ([e][j][6L+3][I]).
template <class T>
int synthesize_state(lbm_plan *p, hipStream_t st) {
  const int E = p->E, L = p->L, EL = E * L;
  const ColMap cm{E, L, 6, 6};
  const int C = cm.ncols();
  OutSlots sl{{0, EL, 2 * EL, 3 * EL, 4 * EL, 5 * EL}};
  CHK((launch_syn<T, SynLevel>(p, p->sa3, p->fin, EL, C, sl, st)));
  OutSlots s2{{6 * EL, 6 * EL + E, 6 * EL + 2 * E, 6 * EL + 3 * E, 6 * EL + 4 * E, 6 * EL + 5 * E}};
  CHK((launch_syn<T, SynSurf>(p, p->sa2, p->fin, E, C, s2, st)));
  CHK(launch_ifft<T>(p, p->fin, p->gin, cm, st));
  return LBM_OK;
}

template <class T>
int get_state_t(lbm_plan *p, double *u, double *v, double *t, double *q, double *ps,
                hipStream_t st) {
  const int J = p->J, I = p->I, L = p->L, E = p->E;
  const size_t n3 = (size_t)E * J * I * L, n2 = (size_t)E * J * I;
  double *dst[4] = {u, v, t, q};
  if (p->have_raw && p->steps_total == 0) {
    for (int f = 0; f < 4; ++f)
      if (dst[f])
        hipLaunchKernelGGL((k_unpack_grid<T>), dim3(nblk(n3)), dim3(256), 0, st, (const T *)p->raw,
                           dst[f], E, J, I, L, 4 * L, f * L, 0, (const T *)nullptr, 0, 0, 1);
    if (ps)
      hipLaunchKernelGGL((k_unpack_grid<T>), dim3(nblk(n2)), dim3(256), 0, st, (const T *)p->psa, ps,
                         E, J, I, 1, 1, 0, 0, (const T *)nullptr, 0, 0, 1);
    HIPCHK(hipGetLastError());
    return LBM_OK;
  }
  CHK(synthesize_state<T>(p, st));
  const int rows = 6 * L + 6;
  const int r0[4] = {0, L, 4 * L, 5 * L};
  for (int f = 0; f < 4; ++f)
    if (dst[f])
      hipLaunchKernelGGL((k_unpack_grid<T>), dim3(nblk(n3)), dim3(256), 0, st, (const T *)p->gin,
                         dst[f], E, J, I, L, rows, r0[f], 0, (const T *)nullptr, 0, 0, 1);
  if (ps) {
    const BRow br{L};
    hipLaunchKernelGGL((k_unpack_grid<T>), dim3(nblk(n2)), dim3(256), 0, st, (const T *)p->gin, ps, E,
                       J, I, 1, rows, 6 * L + 2, p->linear ? 2 : 1, (const T *)p->bsg, br.rows(),
                       br.psbar(), p->nb);
  }
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

// ---------------------------------------------------------------------------
// One time step: synthesis, grid tendencies, analysis, spectral step.
// ---------------------------------------------------------------------------
template <class T>
HistArgs<T> hist_args(const lbm_plan *p) {
  HistArgs<T> h;
  h.mode = p->hist_mode;
  h.K = std::max(p->hist_K, 1);
  h.inistp = p->inistp;
  h.nsteps = p->d_nsteps;
  for (int k = 0; k < 2; ++k) {
    h.spec[k] = (cx<T> *)p->hspec[k];
    h.grid[k] = (T *)p->hgrid[k];
  }
  h.last_spec = p->keep_last ? (cx<T> *)p->hlast_spec : nullptr;
  h.last_grid = p->keep_last ? (T *)p->hlast_grid : nullptr;
  return h;
}

template <class T>
StepArgs<T> step_args(lbm_plan *p, int ecnt) {
  DevTabs<T> &t = p->tabs<T>();
  StepArgs<T> a;
  a.sa3 = (cx<T> *)p->sa3;
  a.sa2 = (cx<T> *)p->sa2;
  a.sb3 = (cx<T> *)p->sb3;
  a.sb2 = (cx<T> *)p->sb2;
  a.tend3 = (const cx<T> *)p->tend3;
  a.tend2 = (const cx<T> *)p->tend2;
  a.kv = t.kv;
  a.kd = t.kd;
  a.ktt = t.ktt;
  a.kq = t.kq;
  a.minv = t.minv;
  a.vinv = t.vinv;
  a.vc = t.vc;
  a.tref = t.tref;
  a.lap = t.lap;
  a.n_of_s = p->d_n_of_s;
  a.ctl = p->ctl;
  a.pend = p->pend;
  a.nsteps = p->d_nsteps;
  a.E = p->E;
  a.M = p->M;
  a.nspec = p->nspec;
  a.nbasic = p->nb;
  a.inistp = p->inistp;
  a.vmask = p->vmask;
  a.skip_low_n = p->skip_low_n;
  a.delt = (T)p->delt;
  a.robert = (T)p->robert;
  a.L = p->L;
  a.ecnt = ecnt;
  a.fixed_set = -1;
  a.fixed_dt = 0;
  return a;
}

// 20 and 40 levels run step kernels specialised for them; other counts
// the wide launchers of lbm_levels.hip (level count at run time).
// k_step's lane mapping: one coefficient and 64 members per wave (UNI,
// member counts that are multiples of 64), the pairs (m, member) of one
// total wavenumber per wave (NG, any other count), both with one basic
// state; with a basic state per member the pairs of one total wavenumber in
// groups of 128 bytes of members (NGM: 8 FP64, 16 FP32; 20 levels and
// member counts that are multiples of the group; measured: 64 members FP64
// 0.920 -> 0.895 ms per T42L20 step, groups of 8 FP32 members lost; at
// T127L40 the FP64 step took 2.36 ms against 1.97 with consecutive pairs);
// else consecutive (coefficient, member) pairs.  NG and NGM buffer offsets
// span all coefficients and must stay below 2^32 bytes.
enum StepMap { STEP_LANES, STEP_UNI, STEP_NG, STEP_NGM };
inline StepMap step_map(const lbm_plan *p, int ecnt) {
  const size_t cs = p->prec == LBM_FP64 ? 16 : 8;
  const size_t span = (size_t)p->nspec * 4 * p->E * p->L * cs;
  const bool fits = span < ((size_t)1 << 32);
  const int g = (int)(128 / cs);
  if (p->nb != 1) return p->L == 20 && p->E % g == 0 && ecnt % g == 0 && fits ? STEP_NGM : STEP_LANES;
  if (ecnt % 64 == 0) return STEP_UNI;
  return fits ? STEP_NG : STEP_LANES;
}
// waves of the NG mapping: ceil((n + 1) ecnt / 64) for each n < M
inline unsigned step_ng_waves(int M, int ecnt) {
  unsigned w = 0;
  for (int n = 0; n < M; ++n) w += (unsigned)(((n + 1) * ecnt + 63) / 64);
  return w;
}

template <class T>
int launch_step_kernel(lbm_plan *p, hipStream_t st, int ecnt) {
  const StepArgs<T> a = step_args<T>(p, ecnt);
  const size_t nthr = (size_t)p->nspec * ecnt;
  // two roles per (s, member): see k_step
  const StepMap map = p->L == 20 || p->L == 40 ? step_map(p, ecnt) : STEP_LANES;
  const bool ng = map == STEP_NG || map == STEP_NGM;
  const dim3 grid(ng ? step_ng_waves(p->M, ecnt) : (unsigned)((nthr + 63) / 64), 2);
  if (p->L == 20) {
    if (map == STEP_UNI)
      hipLaunchKernelGGL((k_step<T, 20, true>), grid, dim3(64), 0, st, a);
    else if (map == STEP_NG)
      hipLaunchKernelGGL((k_step<T, 20, false, true, 1>), grid, dim3(64), 0, st, a);
    else if (map == STEP_NGM)
      hipLaunchKernelGGL((k_step<T, 20, false, true, 128 / (2 * sizeof(T))>), grid, dim3(64), 0, st, a);
    else
      hipLaunchKernelGGL((k_step<T, 20, false>), grid, dim3(64), 0, st, a);
  } else if (p->L == 40) {
    if (map == STEP_UNI)
      hipLaunchKernelGGL((k_step<T, 40, true>), grid, dim3(64), 0, st, a);
    else if (map == STEP_NG)
      hipLaunchKernelGGL((k_step<T, 40, false, true, 1>), grid, dim3(64), 0, st, a);
    else
      hipLaunchKernelGGL((k_step<T, 40, false>), grid, dim3(64), 0, st, a);
  } else {
    HIPCHK(launch_step_wide<T>(a, grid, st));
  }
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

// Event k marks the start of profiling slot k (slot order of lbm_profile).
inline int mark(lbm_plan *p, int k, hipStream_t st) {
  if (p->prof) HIPCHK(hipEventRecord(p->ev[k], st));
  return LBM_OK;
}

// Arguments of k_tend for the plan's buffers.
template <class T>
TendArgs<T> tend_args(lbm_plan *p, bool raw_path, const ActiveMask &am) {
  DevTabs<T> &t = p->tabs<T>();
  TendArgs<T> ta;
  ta.gin = (const T *)p->gin;
  ta.raw = (raw_path && p->have_raw) ? (const T *)p->raw : nullptr;
  ta.bsg = (const T *)p->bsg;
  ta.frc = (const T *)p->frc;
  ta.gout = (T *)p->gout;
  ta.psa = (T *)p->psa;
  ta.psb = (T *)p->psb;
  ta.pend = p->pend;
  ta.vc = t.vc;
  ta.tref = t.tref;
  ta.rcos = t.rcos;
  ta.ctl = p->ctl;
  ta.E = p->E;
  ta.J = p->J;
  ta.I = p->I;
  ta.L = p->L;
  ta.nbasic = p->nb;
  ta.nfrc = p->nfrc;
  ta.strictly_linear = p->strictly_linear || p->linear;
  ta.linear = p->linear;
  ta.robert = (T)p->robert;
  ta.am = am;
  ta.h = hist_args<T>(p);
  ta.traj = (T *)p->traj_cur;
  ta.fout = (cx<T> *)p->fout;
  ta.C = ColMap{p->E, p->L, 9, 2}.ncols();
  ta.M = p->M;
  ta.ne = p->E;
  ta.fin = (const cx<T> *)p->fin;
  return ta;
}

// The grid phase (tuning choice 4): the inverse FFT, k_tend and the forward
// FFT as three launches (0); the forward FFT inside k_tend (1 and 2: two
// group sizes, 128 longitudes); all three in k_grid (3 and 4:
// grid_members<T, NI>(v) members per block, 128 or 384 longitudes with
// nmax + 1 <= grid_pf_m; the built-in choice there), except on the raw
// first step after set_initial_perturbation, which takes choice 0.
inline bool tend_fused(const lbm_plan *p) {
  return p->I == 128 && (p->tune[4] == 1 || p->tune[4] == 2);
}
// k_grid applies at 128 longitudes with nmax + 1 <= grid_pf_m.  (Its
// 384-longitude instantiation is switched off: its first GPU run faulted
// with a memory aperture violation whose cause is not found.)
inline bool grid_fits(const lbm_plan *p) { return p->I == 128 && p->M <= grid_pf_m<128>(); }
inline bool grid_fused(const lbm_plan *p, bool raw) { return grid_fits(p) && p->tune[4] >= 3 && !raw; }
// Built-in grid phase of a plan of E members: k_grid where it applies and
// the ensemble gives it enough blocks, else the three launches.  k_grid
// runs one block per latitude and member group, each through every level
// in turn, so with a few member groups per latitude the GPU is not filled
// and the three launches finish first.  Measured at T42L20 (ms per step,
// three launches / k_grid): FP64 1 member 0.174 / 0.217, 4: 0.192 / 0.216,
// 8: 0.257 / 0.238 (4 per block); FP32 4 members 0.160 / 0.177 (4 per
// block), 8: 0.189 / 0.189, 16: 0.261 / 0.223 (4 per block) and 0.247 (8
// per block), 24: 0.318 / 0.306 (4) and 0.265 (8).
inline int builtin_grid_choice(const lbm_plan *p) {
  if (!grid_fits(p)) return 0;
  if (p->prec == LBM_FP64) return p->E > 4 ? 4 : 0;
  return p->E > 16 ? 4 : p->E > 8 ? 3 : 0;
}

template <class T, int NI, int MB>
int launch_grid_mb(lbm_plan *p, const TendArgs<T> &ta, int ne, hipStream_t st) {
  // JP x NSIB blocks per (8-member tile, JP latitudes), 8 apart: see k_grid
  constexpr int NSIB = 8 / MB, JP = sizeof(T) == 4 ? 2 : 1;
  const long units = (long)((ne + 7) / 8) * (p->J / JP);
  TendArgs<T> tf = ta;
  tf.ne = ne;
  const dim3 grid((unsigned)((units + 7) / 8 * 8 * NSIB * JP)), block(grid_threads<NI>(MB));
  if (p->nb == 1)
    hipLaunchKernelGGL((k_grid<T, NI, MB, true>), grid, block, 0, st, tf);
  else
    hipLaunchKernelGGL((k_grid<T, NI, MB, false>), grid, block, 0, st, tf);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

// (128 longitudes: grid_fits)
template <class T>
int launch_grid(lbm_plan *p, const TendArgs<T> &ta, int ne, hipStream_t st) {
  if (p->tune[4] == 3) return launch_grid_mb<T, 128, grid_members<T, 128>(3)>(p, ta, ne, st);
  return launch_grid_mb<T, 128, grid_members<T, 128>(4)>(p, ta, ne, st);
}

// The inverse FFT of the synthesised columns unless k_grid does it; the
// profiling marks 2 and 3 bracket it.
template <class T>
int launch_grid_in(lbm_plan *p, const TendArgs<T> &ta, const ActiveMask &am, hipStream_t st) {
  CHK(mark(p, 2, st));
  if (!grid_fused(p, ta.raw != nullptr))
    CHK(launch_ifft<T>(p, p->fin, p->gin, ColMap{p->E, p->L, 6, 6}, st, 0, am));
  return LBM_OK;
}

// k_tend for ne members, then (unfused) the forward FFT of its rows; or
// k_grid.  The profiling marks 3 and 4 bracket k_tend or k_grid.
template <class T>
int launch_tend(lbm_plan *p, const TendArgs<T> &ta, int ne, hipStream_t st) {
  const ColMap cmo{p->E, p->L, 9, 2};
  CHK(mark(p, 3, st));
  if (grid_fused(p, ta.raw != nullptr)) {
    CHK(launch_grid<T>(p, ta, ne, st));
    CHK(mark(p, 4, st));
    return LBM_OK;
  }
  if (tend_fused(p)) {
    // blocks of (MB members, latitude), MB x 128 threads
    TendArgs<T> tf = ta;
    tf.ne = ne;
    if (p->tune[4] == 1) {
      constexpr int MB = tend_fft_members<T>(1);
      hipLaunchKernelGGL((k_tend<T, true, MB>), dim3((unsigned)((ne + MB - 1) / MB * p->J)), dim3(MB * 128),
                         0, st, tf);
    } else {
      constexpr int MB = tend_fft_members<T>(2);
      hipLaunchKernelGGL((k_tend<T, true, MB>), dim3((unsigned)((ne + MB - 1) / MB * p->J)), dim3(MB * 128),
                         0, st, tf);
    }
  } else {
    hipLaunchKernelGGL((k_tend<T, false>), dim3(ne * p->J * (p->I / grid_chunk(p->I))),
                       dim3(grid_chunk(p->I)), 0, st, ta);
  }
  HIPCHK(hipGetLastError());
  CHK(mark(p, 4, st));
  if (!tend_fused(p)) CHK(launch_fft<T>(p, p->gout, p->fout, cmo, st, 0, ta.am));
  return LBM_OK;
}

// One time step: synthesis, inverse FFT, grid tendencies, forward FFT,
// analysis, spectral step.  raw_path: the first step after a raw initial
// state (k_tend reads the raw grid).
// ecnt: leading members covered by the launches, or 0 when every member is
// live through the step (no activity mask)
// tend_only: stop after the analysed tendencies (tend3, tend2) of the
// current state, with no history sums and no spectral step.
template <class T>
int one_step(lbm_plan *p, hipStream_t st, bool raw_path, int ecnt, bool tend_only = false) {
  const int E = p->E, L = p->L, EL = E * L;
  DevTabs<T> &t = p->tabs<T>();
  const ColMap cmi{E, L, 6, 6};
  const int Ci = cmi.ncols();
  const ActiveMask am = loop_mask(p, true, ecnt);
  const int ne = ecnt > 0 ? ecnt : E;  // members launched
  // The surface synthesis, the MJO forcing of this call, the history sums of
  // the top-of-step state and the surface analysis run on a second stream
  // beside the level kernels (none reads what the other side writes), each
  // joined before its first consumer: the inverse FFT, k_tend, k_step and
  // the steady-forcing / step kernels.  Profiling keeps one stream, so its
  // per-slot events time each launch; so do launches of at most 8 members,
  // where the fork and join cost more than the overlap gains (measured at
  // T42L20: 1 member 0.175 -> 0.171 ms per step FP64, 2 members 0.147 ->
  // 0.140 FP32; 16 members lose without the fork).
  const bool fork = !p->prof && p->side != nullptr && ne > 8;
  hipStream_t sd = fork ? p->side : st;
  enum { F_FFT, F_SURF, F_MJO, F_HIST, F_ANAS, F_START };
  auto fork_at = [&](int k) -> int {
    if (fork) {
      HIPCHK(hipEventRecord(p->fj[k], st));
      HIPCHK(hipStreamWaitEvent(sd, p->fj[k], 0));
    }
    return LBM_OK;
  };
  auto side_done = [&](int k) -> int {
    if (fork) HIPCHK(hipEventRecord(p->fj[k], sd));
    return LBM_OK;
  };
  auto join = [&](int k) -> int {
    if (fork) HIPCHK(hipStreamWaitEvent(st, p->fj[k], 0));
    return LBM_OK;
  };
  const bool hist = !tend_only && (p->hist_mode != 0 || p->keep_last);
  auto launch_mjo = [&]() -> int {
    // this call's MJO temperature forcing (forcing.py:113-156)
    const size_t n = (size_t)p->J * L * p->I;
    hipLaunchKernelGGL((k_mjo<T>), dim3(nblk((size_t)p->nfrc * n)), dim3(256), 0, sd,
                       (const T *)p->mjo, (T *)p->frc, (const Ctl *)p->ctl, p->nfrc, p->mjo_nd,
                       p->mjo_nt, p->mjo_spin, p->mjo_call0, p->mjo_single, n);
    HIPCHK(hipGetLastError());
    return side_done(F_MJO);
  };
  auto launch_hist = [&]() -> int {
    const size_t nthr = (size_t)p->nspec * ne;
    hipLaunchKernelGGL((k_hist_acc<T>), dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, sd,
                       (const cx<T> *)p->sa3, hist_args<T>(p), (const Ctl *)p->ctl, E, L, p->nspec,
                       ne);
    HIPCHK(hipGetLastError());
    return side_done(F_HIST);
  };
  CHK(fork_at(F_START));
  {
    OutSlots sl{{0, EL, 2 * EL, 3 * EL, 4 * EL, 5 * EL}};
    CHK(mark(p, 0, st));
    CHK((launch_syn<T, SynLevel>(p, p->sa3, p->fin, EL, Ci, sl, st, SynLevel::NIN, nullptr, nullptr,
                                 nullptr, am)));
    OutSlots s2{{6 * EL, 6 * EL + E, 6 * EL + 2 * E, 6 * EL + 3 * E, 6 * EL + 4 * E, 6 * EL + 5 * E}};
    CHK(mark(p, 1, st));
    CHK((launch_syn<T, SynSurf>(p, p->sa2, p->fin, E, Ci, s2, sd, SynSurf::NIN, nullptr, nullptr,
                                nullptr, am)));
    CHK(side_done(F_SURF));
    if (fork && p->mjo) CHK(launch_mjo());
    if (fork && hist) CHK(launch_hist());
  }
  const TendArgs<T> ta = tend_args<T>(p, raw_path, am);
  const ColMap cmo{E, L, 9, 2};
  const int Co = cmo.ncols();
  CHK(join(F_SURF));
  CHK(launch_grid_in<T>(p, ta, am, st));
  if (!fork && p->mjo) CHK(launch_mjo());
  if (p->mjo) CHK(join(F_MJO));
  CHK(launch_tend<T>(p, ta, ne, st));
  CHK(fork_at(F_FFT));
  CHK(mark(p, 5, st));
  {
    // (FU, FV, KE) -> vor, div; (XU, XV, GT) -> t; (YU, YV, G) -> q
    InSlots vd{{0, EL, 2 * EL}}, ft{{3 * EL, 4 * EL, 5 * EL}}, fq{{6 * EL, 7 * EL, 8 * EL}};
    CHK((launch_ana<T, AnaTendVD>(p, p->fout, p->tend3, EL, Co, vd, st, 4, 0, nullptr, nullptr,
                                  nullptr, am)));
    // temperature and humidity beside it on the second stream (their own
    // output slots), the surface analysis after it
    CHK((launch_ana<T, AnaFluxDiv>(p, p->fout, p->tend3, EL, Co, ft, sd, 4, 2, nullptr, nullptr,
                                   nullptr, am)));
    CHK((launch_ana<T, AnaFluxDiv>(p, p->fout, p->tend3, EL, Co, fq, sd, 4, 3, nullptr, nullptr,
                                   nullptr, am)));
  }
  CHK(side_done(F_ANAS));
  InSlots s2{{9 * EL}};
  CHK(mark(p, 6, st));
  CHK((launch_ana<T, AnaScalar>(p, p->fout, p->tend2, 2 * E, Co, s2, st, 1, 0, nullptr, nullptr,
                                nullptr, am)));
  CHK(join(F_ANAS));
  if (p->sfv || p->sfp) {
    const size_t nthr = (size_t)p->nspec * ne;
    hipLaunchKernelGGL((k_add_steady<T>), dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, st,
                       (cx<T> *)p->tend3, (cx<T> *)p->tend2, (const cx<T> *)p->sfv,
                       (const cx<T> *)p->sfp, E, L, p->nspec, p->nsf, ne);
    HIPCHK(hipGetLastError());
  }
  if (tend_only) return LBM_OK;
  if (!fork && hist) CHK(launch_hist());
  if (hist) CHK(join(F_HIST));
  CHK(mark(p, 7, st));
  CHK(launch_step_kernel<T>(p, st, ne));
  CHK(mark(p, 8, st));
  if (p->resp_set)
    hipLaunchKernelGGL((k_response<T>), dim3(ne), dim3(256), 0, st, (const cx<T> *)p->sa3,
                       (const cx<T> *)p->obs, (const T *)p->wlev, (const T *)p->wstep, p->nwstep, E,
                       L, p->nspec, (const Ctl *)p->ctl, (const int *)p->d_nsteps, p->inistp,
                       p->jacc);
  hipLaunchKernelGGL(k_tick, dim3(1), dim3(1), 0, st, p->ctl);
  HIPCHK(hipGetLastError());
  if (p->prof) {
    HIPCHK(hipEventSynchronize(p->ev[lbm_plan::NPROF]));
    for (int k = 0; k < lbm_plan::NPROF; ++k) {
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, p->ev[k], p->ev[k + 1]));
      p->prof_ms[k] += ms;
      p->prof_n[k] += 1;
    }
  }
  return LBM_OK;
}

int step_dispatch(lbm_plan *p, hipStream_t st, int ecnt, bool raw_path = false) {
  return p->prec == LBM_FP64 ? one_step<double>(p, st, raw_path, ecnt)
                             : one_step<float>(p, st, raw_path, ecnt);
}

// the graph of nsteps steps (kind 0: the 8-step chunk, 1: one step) over
// the first ecnt members, captured on first use
int capture_graph(lbm_plan *p, int kind, int nsteps, int ecnt, hipStream_t st,
                  hipGraphExec_t *out) {
  if (p->prof) return fail(LBM_ERR_STATE, "profiling runs eager launches only (use_graph=0)");
  const auto key = std::make_pair(kind, ecnt);
  auto it = p->graphs.find(key);
  if (it != p->graphs.end()) {
    *out = it->second;
    return LBM_OK;
  }
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int rc = LBM_OK;
  for (int s = 0; s < nsteps && rc == LBM_OK; ++s) rc = step_dispatch(p, st, ecnt);
  hipError_t e2 = hipStreamEndCapture(st, &g);
  if (rc != LBM_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e2 != hipSuccess) return fail(LBM_ERR_RUNTIME, "graph capture failed: %s", hipGetErrorString(e2));
  hipGraphExec_t ge = nullptr;
  hipError_t e3 = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e3 != hipSuccess) return fail(LBM_ERR_RUNTIME, "graph instantiate failed: %s", hipGetErrorString(e3));
  p->graphs[key] = ge;
  *out = ge;
  return LBM_OK;
}

int check_bad(lbm_plan *p, hipStream_t st, long long base_step) {
  Ctl h;
  HIPCHK(hipMemcpyAsync(&h, p->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (h.bad_step != INT_MAX)
    return fail(LBM_ERR_NONFINITE, "model blew up at step %d (t=%.2f d)", h.bad_step,
                (double)std::max(0LL, (long long)h.bad_step - p->inistp) * p->delt / 86400.0);
  (void)base_step;
  return LBM_OK;
}

template <class T>
int final_check(lbm_plan *p, hipStream_t st) {
  const size_t n3 = (size_t)p->nspec * 4 * p->E * p->L, n2 = (size_t)p->nspec * 3 * p->E;
  hipLaunchKernelGGL((k_check_spec<T>), dim3(256), dim3(256), 0, st, (const cx<T> *)p->sa3, n3,
                     (const cx<T> *)p->sa2, n2, p->ctl);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

// ---------------------------------------------------------------------------
// Autotuner: each tunable launch of the time step is timed in its variants on
// the plan's own buffers and shapes (HIP events on the plan's stream, the
// best of three trials of `reps` launches), and the fastest is kept.  The
// variants differ only in block shape and register budget, never in
// arithmetic, so results are bit-identical whichever is chosen.
// ---------------------------------------------------------------------------
size_t state_bytes(const lbm_plan *p);
int copy_state(lbm_plan *p, void *buf, bool save, hipStream_t st);

// Tuning choice 4: the grid phase as three launches (0), with the forward
// FFT inside k_tend (1, 2; 128 longitudes) or in one k_grid launch (3, 4;
// where grid_fits), timed on the plan's buffers with no history sums or
// trajectory; the state the tendencies update (deferred ps' updates, the
// blow-up flag) is saved before and restored after.
template <class T, class Timed>
int autotune_tend_fft(lbm_plan *p, Timed &timed, char *buf, size_t len) {
  constexpr int NV = 5;
  bool on[NV] = {true, p->I == 128, p->I == 128, grid_fits(p), grid_fits(p)};
  if (!on[1] && !on[3]) {
    p->tune[4] = 0;
    std::snprintf(buf, len, "; tend_fft: separate (no fused variant at %d longitudes)", p->I);
    return LBM_OK;
  }
  hipStream_t st = p->own;
  void *save = nullptr;
  HIPCHK(hipMalloc(&save, state_bytes(p)));
  int rc = copy_state(p, save, true, st);
  TendArgs<T> ta = tend_args<T>(p, false, loop_mask(p, true, 0));
  ta.h.mode = 0;
  ta.h.last_grid = nullptr;
  ta.h.last_spec = nullptr;
  ta.traj = nullptr;
  const ActiveMask am = loop_mask(p, true, 0);
  double ms[NV] = {0, 0, 0, 0, 0};
  for (int v = 0; v < NV && rc == LBM_OK; ++v) {
    if (!on[v]) continue;
    p->tune[4] = v;
    rc = timed([&] {
      CHK(launch_grid_in<T>(p, ta, am, st));
      return launch_tend<T>(p, ta, p->E, st);
    }, &ms[v]);
  }
  if (copy_state(p, save, false, st) != LBM_OK && rc == LBM_OK) rc = LBM_ERR_RUNTIME;
  if (hipStreamSynchronize(st) != hipSuccess && rc == LBM_OK) rc = fail(LBM_ERR_RUNTIME, "sync failed");
  (void)hipFree(save);
  if (rc != LBM_OK) return rc;
  int best = 0;
  for (int v = 1; v < NV; ++v)
    if (on[v] && ms[v] < ms[best]) best = v;
  p->tune[4] = best;
  const char *lab[NV] = {"separate", "tend+fft", "tend+fft", "grid", "grid"};
  const bool i128 = p->I == 128;
  const int mb[NV] = {0, tend_fft_members<T>(1), tend_fft_members<T>(2),
                      i128 ? grid_members<T, 128>(3) : 0, i128 ? grid_members<T, 128>(4) : 0};
  std::string r = "; tend_fft: separate=" + std::to_string(ms[0]).substr(0, 6);
  for (int v = 1; v < NV; ++v) {
    if (!on[v]) continue;
    char one[64];
    std::snprintf(one, sizeof(one), " %s/%d=%.4f", lab[v], mb[v], ms[v]);
    r += one;
  }
  std::snprintf(buf, len, "%s ms -> %d (%s)", r.c_str(), best, lab[best]);
  return LBM_OK;
}

template <class T>
int autotune_t(lbm_plan *p, int reps, std::string &rep) {
  const int E = p->E, L = p->L, EL = E * L;
  hipStream_t st = p->own;
  const ColMap cmi{E, L, 6, 6}, cmo{E, L, 9, 2};
  const int Ci = cmi.ncols(), Co = cmo.ncols();
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  auto timed = [&](auto launch, double *ms) -> int {
    CHK(launch());  // warm (first launch of a variant loads its code object)
    double best = 1e300;
    for (int trial = 0; trial < 3; ++trial) {
      HIPCHK(hipEventRecord(e0, st));
      for (int r = 0; r < reps; ++r) CHK(launch());
      HIPCHK(hipEventRecord(e1, st));
      HIPCHK(hipEventSynchronize(e1));
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, e0, e1));
      best = std::min(best, (double)t / reps);
    }
    *ms = best;
    return LBM_OK;
  };
  const char *names[lbm_plan::NTUNE] = {"ifft_rows", "fft_rows", "syn_waves", "ana_waves", "tend_fft"};
  char buf[256];
  for (int knob = 0; knob < lbm_plan::NTUNE; ++knob) {
    if (knob == 4) {
      CHK(autotune_tend_fft<T>(p, timed, buf, sizeof(buf)));
      rep += buf;
      continue;
    }
    if (knob < 2 && p->ffac.np > 0) {
      // one Stockham shape per length: nothing to choose
      p->tune[knob] = 0;
      std::snprintf(buf, sizeof(buf), "%s%s: stockham %d", knob ? "; " : "", names[knob],
                    fft_any_rows(p->ffac.NC));
      rep += buf;
      continue;
    }
    double ms[3];
    for (int v = 0; v < 3; ++v) {
      p->tune[knob] = v;
      int rc = LBM_OK;
      if (knob == 0)
        rc = timed([&] { return launch_ifft<T>(p, p->fin, p->gin, cmi, st); }, &ms[v]);
      else if (knob == 1)
        rc = timed([&] { return launch_fft<T>(p, p->gout, p->fout, cmo, st); }, &ms[v]);
      else if (knob == 2)
        rc = timed([&] {
          OutSlots sl{{0, EL, 2 * EL, 3 * EL, 4 * EL, 5 * EL}};
          return launch_syn<T, SynLevel>(p, p->sa3, p->fin, EL, Ci, sl, st);
        }, &ms[v]);
      else
        rc = timed([&] {
          InSlots vd{{0, EL, 2 * EL}}, ft{{3 * EL, 4 * EL, 5 * EL}}, fq{{6 * EL, 7 * EL, 8 * EL}};
          CHK((launch_ana<T, AnaTendVD>(p, p->fout, p->tend3, EL, Co, vd, st, 4, 0)));
          CHK((launch_ana<T, AnaFluxDiv>(p, p->fout, p->tend3, EL, Co, ft, st, 4, 2)));
          return launch_ana<T, AnaFluxDiv>(p, p->fout, p->tend3, EL, Co, fq, st, 4, 3);
        }, &ms[v]);
      if (rc != LBM_OK) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return rc;
      }
    }
    int best = 0;
    for (int v = 1; v < 3; ++v)
      if (ms[v] < ms[best]) best = v;
    p->tune[knob] = best;
    auto label = [&](int v) -> int {
      if (knob == 2) return SYN_OCC[v];
      if (knob == 3) return ANA_OCC[v];
      const bool inv = knob == 0;
      if (v == 0)
        return p->I == 128 ? (inv ? FftShape<T, 128, true>::ROWS : FftShape<T, 128>::ROWS)
                           : (inv ? FftShape<T, 384, true>::ROWS : FftShape<T, 384>::ROWS);
      return p->I == 128 ? (inv ? fft_alt_rows<T, 128, true>(v) : fft_alt_rows<T, 128, false>(v))
                         : (inv ? fft_alt_rows<T, 384, true>(v) : fft_alt_rows<T, 384, false>(v));
    };
    std::snprintf(buf, sizeof(buf), "%s%s: %d=%.4f %d=%.4f %d=%.4f ms -> %d", knob ? "; " : "",
                  names[knob], label(0), ms[0], label(1), ms[1], label(2), ms[2], label(best));
    rep += buf;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return LBM_OK;
}

// ---------------------------------------------------------------------------
// Stand-alone tendencies and step (Dynamics.tendencies, Integrator.step) on
// caller buffers, the previous time level on the grid and the partial state
// update of set_initial_perturbation.
// ---------------------------------------------------------------------------
int ensure_aux(lbm_plan *p) {
  if (p->x_sa3) return LBM_OK;
  const size_t rs = p->rsize(), cs = 2 * rs, E = p->E, L = p->L, J = p->J, I = p->I;
  const size_t EL = E * L, ns = p->nspec;
  struct A {
    void **ptr;
    size_t bytes;
  } al[] = {{&p->x_sa3, cs * ns * 4 * EL}, {&p->x_sb3, cs * ns * 4 * EL},
            {&p->x_sa2, cs * ns * 3 * E},  {&p->x_psa, rs * E * J * I},
            {&p->x_psb, rs * E * J * I},   {&p->x_raw, rs * E * J * 4 * L * I},
            {(void **)&p->x_pend, sizeof(int) * E}, {(void **)&p->x_ctl, sizeof(Ctl)}};
  for (auto &a : al) HIPCHK(hipMalloc(a.ptr, std::max<size_t>(a.bytes, 16)));
  if (p->linear) HIPCHK(hipMalloc(&p->x_sb2, cs * ns * 3 * E));
  return LBM_OK;
}

// exchanges the run's prognostic buffers with the second copy
void swap_state(lbm_plan *p) {
  std::swap(p->sa3, p->x_sa3);
  std::swap(p->sb3, p->x_sb3);
  std::swap(p->sa2, p->x_sa2);
  std::swap(p->sb2, p->x_sb2);
  std::swap(p->psa, p->x_psa);
  std::swap(p->psb, p->x_psb);
  std::swap(p->raw, p->x_raw);
  std::swap(p->pend, p->x_pend);
  std::swap(p->ctl, p->x_ctl);
}

// (vor, div, t, q) of a [s][nf3][L][E] array and pi (field f2 of a
// [s][nf2][E] array) to the reference's spectral layout; null outputs skipped
template <class T>
int unpack_spec5(lbm_plan *p, const void *s3, int nf3, const void *s2, int nf2, int f2,
                 double *const out[5], hipStream_t st) {
  const int E = p->E, M = p->M, L = p->L;
  const size_t n3 = (size_t)E * M * M * L, n2 = (size_t)E * M * M;
  for (int f = 0; f < 4; ++f)
    if (out[f])
      hipLaunchKernelGGL((k_unpack_spec<T>), dim3(nblk(n3)), dim3(256), 0, st, (const cx<T> *)s3,
                         out[f], E, M, L, nf3, f, E * L);
  if (out[4])
    hipLaunchKernelGGL((k_unpack_spec<T>), dim3(nblk(n2)), dim3(256), 0, st, (const cx<T> *)s2,
                       out[4], E, M, 1, nf2, f2, E);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

template <class T>
int pack_spec5(lbm_plan *p, const double *const in[5], void *s3, int nf3, void *s2, int nf2, int f2,
               hipStream_t st) {
  const int E = p->E, M = p->M, L = p->L;
  const size_t n3 = (size_t)p->nspec * E * L, n2 = (size_t)p->nspec * E;
  for (int f = 0; f < 4; ++f)
    hipLaunchKernelGGL((k_pack_spec<T>), dim3(nblk(n3)), dim3(256), 0, st, in[f], (cx<T> *)s3, E, M,
                       L, nf3, f, E * L);
  hipLaunchKernelGGL((k_pack_spec<T>), dim3(nblk(n2)), dim3(256), 0, st, in[4], (cx<T> *)s2, E, M, 1,
                     nf2, f2, E);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

// Dynamics.tendencies (dynamics.py:156-271) of a caller grid state with the
// plan's forcings: the same set-up and kernels as the first step after a
// raw initial perturbation, on the second copy of the state.
template <class T>
int tendencies_t(lbm_plan *p, const double *u, const double *v, const double *t, const double *q,
                 const double *ps, double *const out[5], hipStream_t st) {
  CHK(ensure_aux(p));
  const bool have_raw = p->have_raw, raw_pending = p->raw_pending, keep_last = p->keep_last;
  const int hist_mode = p->hist_mode;
  swap_state(p);
  p->hist_mode = 0;
  p->keep_last = false;
  int rc = set_state_t<T>(p, u, v, t, q, ps, st);
  Ctl c0{0, p->have_raw ? 1 : 0, INT_MAX, 0};
  if (rc == LBM_OK && (hipMemcpyAsync(p->ctl, &c0, sizeof(Ctl), hipMemcpyHostToDevice, st) != hipSuccess ||
                       hipStreamSynchronize(st) != hipSuccess))
    rc = fail(LBM_ERR_RUNTIME, "control block upload failed");
  if (rc == LBM_OK) rc = one_step<T>(p, st, true, 0, true);
  if (rc == LBM_OK) rc = unpack_spec5<T>(p, p->tend3, 4, p->tend2, 2, 0, out, st);
  swap_state(p);
  p->have_raw = have_raw;
  p->raw_pending = raw_pending;
  p->hist_mode = hist_mode;
  p->keep_last = keep_last;
  return rc;
}

// matrix set of a time step: a start-up dt, delt or the free dt; -1 if none
int find_set(const lbm_plan *p, double dt) {
  for (int set = 0; set < p->nset; ++set)
    if (std::fabs(step_dt(p, set) - dt) <= 1e-12 * std::fabs(dt)) return set;
  return -1;
}

// SemiImplicit.prepare / VerticalDiffusion.prepare (integrator.py:130-163,
// 216-228) for the free matrix set
int prepare_dt(lbm_plan *p, double dt) {
  if (!(dt > 0.0) || !std::isfinite(dt)) return fail(LBM_ERR_VALUE, "dt must be positive, got %g", dt);
  if (!p->basic_set) return fail(LBM_ERR_STATE, "set the basic state before preparing a time step");
  p->user_dt = dt;
  CHK(build_matrices(p, p->nset - 1, p->nset));
  CHK(build_vdiff(p, p->nset - 1, p->nset));
  CHK(upload_operators<double>(p));
  if (p->prec != LBM_FP64) CHK(upload_operators<float>(p));
  return LBM_OK;
}

// Integrator.step (integrator.py:264-315) on caller buffers: k_step with a
// fixed dt on the second copy of the spectral state
template <class T>
int step_t(lbm_plan *p, double dt, const double *const prev[5], const double *const tend[5],
           double *const out[5], hipStream_t st) {
  CHK(ensure_aux(p));
  int set = find_set(p, dt);
  if (set < 0) {
    CHK(prepare_dt(p, dt));
    set = p->nset - 1;
  }
  // previous level: (vor, div, t, q) as the filtered level, pi where the step
  // reads pi at t-dt (tendency field 1); tendencies in their usual slots
  CHK(pack_spec5<T>(p, prev, p->x_sb3, 4, p->tend2, 2, 1, st));
  CHK(pack_spec5<T>(p, tend, p->tend3, 4, p->tend2, 2, 0, st));
  StepArgs<T> a = step_args<T>(p, p->E);
  a.sa3 = (cx<T> *)p->x_sa3;
  a.sb3 = (cx<T> *)p->x_sb3;
  a.sa2 = (cx<T> *)p->x_sa2;
  a.sb2 = nullptr;
  a.pend = p->x_pend;
  a.fixed_set = set;
  a.fixed_dt = (T)dt;
  const dim3 grid((unsigned)(((size_t)p->nspec * p->E + 63) / 64), 2);
  if (p->L == 20)
    hipLaunchKernelGGL((k_step<T, 20, false>), grid, dim3(64), 0, st, a);
  else if (p->L == 40)
    hipLaunchKernelGGL((k_step<T, 40, false>), grid, dim3(64), 0, st, a);
  else
    HIPCHK(launch_step_wide<T>(a, grid, st));
  HIPCHK(hipGetLastError());
  return unpack_spec5<T>(p, p->x_sa3, 4, p->x_sa2, 3, 0, out, st);
}

// The deferred ps' update of the last step applied now (the next step then
// reads psa / psb as they are).
template <class T>
int settle_ps(lbm_plan *p, hipStream_t st) {
  if (p->linear || p->steps_total == 0) return LBM_OK;
  CHK(synthesize_state<T>(p, st));
  const int E = p->E, J = p->J, I = p->I, L = p->L;
  const BRow br{L};
  hipLaunchKernelGGL((k_settle_ps<T>), dim3(nblk((size_t)E * J * I)), dim3(256), 0, st,
                     (const T *)p->gin, 6 * L + 6, 6 * L + 2, (const T *)p->bsg, br.rows(), br.psbar(),
                     (T *)p->psa, (T *)p->psb, (const int *)p->pend, (T)p->robert, E, J, I, p->nb);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemsetAsync(p->pend, 0, E * sizeof(int), st));
  return LBM_OK;
}

// grid fields of time level 0 (ga) or 1 (gb, the filtered previous level)
template <class T>
int get_level_t(lbm_plan *p, int level, double *u, double *v, double *t, double *q, double *ps,
                hipStream_t st) {
  if (level == 0 || (p->steps_total == 0 && p->levels_equal))
    return get_state_t<T>(p, u, v, t, q, ps, st);
  const int J = p->J, I = p->I, L = p->L, E = p->E;
  if (p->steps_total == 0 && p->rawb) {
    // the previous level as the partial update left it: its raw grid and ps'
    const size_t n3 = (size_t)E * J * I * L, n2 = (size_t)E * J * I;
    double *dst[4] = {u, v, t, q};
    for (int f = 0; f < 4; ++f)
      if (dst[f])
        hipLaunchKernelGGL((k_unpack_grid<T>), dim3(nblk(n3)), dim3(256), 0, st, (const T *)p->rawb,
                           dst[f], E, J, I, L, 4 * L, f * L, 0, (const T *)nullptr, 0, 0, 1);
    if (ps)
      hipLaunchKernelGGL((k_unpack_grid<T>), dim3(nblk(n2)), dim3(256), 0, st, (const T *)p->psb, ps,
                         E, J, I, 1, 1, 0, 0, (const T *)nullptr, 0, 0, 1);
    HIPCHK(hipGetLastError());
    return LBM_OK;
  }
  CHK(settle_ps<T>(p, st));
  std::swap(p->sa3, p->sb3);
  if (p->linear) std::swap(p->sa2, p->sb2);
  const int rc = synthesize_state<T>(p, st);
  std::swap(p->sa3, p->sb3);
  if (p->linear) std::swap(p->sa2, p->sb2);
  CHK(rc);
  const size_t n3 = (size_t)E * J * I * L, n2 = (size_t)E * J * I;
  const int rows = 6 * L + 6;
  const int r0[4] = {0, L, 4 * L, 5 * L};
  double *dst[4] = {u, v, t, q};
  for (int f = 0; f < 4; ++f)
    if (dst[f])
      hipLaunchKernelGGL((k_unpack_grid<T>), dim3(nblk(n3)), dim3(256), 0, st, (const T *)p->gin,
                         dst[f], E, J, I, L, rows, r0[f], 0, (const T *)nullptr, 0, 0, 1);
  if (ps) {
    const BRow br{L};
    if (p->linear)
      hipLaunchKernelGGL((k_unpack_grid<T>), dim3(nblk(n2)), dim3(256), 0, st, (const T *)p->gin, ps,
                         E, J, I, 1, rows, 6 * L + 2, 2, (const T *)p->bsg, br.rows(), br.psbar(), p->nb);
    else
      hipLaunchKernelGGL((k_unpack_grid<T>), dim3(nblk(n2)), dim3(256), 0, st, (const T *)p->psb, ps,
                         E, J, I, 1, 1, 0, 0, (const T *)nullptr, 0, 0, 1);
  }
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

// set_initial_perturbation (model.py:142-148): the given fields replace both
// time levels, the others keep each level's values.
template <class T>
int set_state_keep_t(lbm_plan *p, const double *const in[5], hipStream_t st) {
  const int J = p->J, I = p->I, L = p->L, E = p->E;
  const size_t n3 = (size_t)E * J * I * L, n2 = (size_t)E * J * I;
  const bool two = p->steps_total > 0 || !p->levels_equal;
  double *buf = nullptr;
  HIPCHK(hipMalloc(&buf, (two ? 2 : 1) * (4 * n3 + n2) * sizeof(double)));
  auto level_ptrs = [&](double *base, double *o[5]) {
    for (int f = 0; f < 4; ++f) o[f] = base + (size_t)f * n3;
    o[4] = base + 4 * n3;
  };
  double *ga[5], *gb[5];
  level_ptrs(buf, ga);
  if (two) level_ptrs(buf + 4 * n3 + n2, gb);
  int rc = get_level_t<T>(p, 0, ga[0], ga[1], ga[2], ga[3], ga[4], st);
  if (rc == LBM_OK && two) rc = get_level_t<T>(p, 1, gb[0], gb[1], gb[2], gb[3], gb[4], st);
  const double *a[5], *b[5];
  for (int f = 0; f < 5; ++f) {
    a[f] = in[f] ? in[f] : ga[f];
    b[f] = in[f] ? in[f] : (two ? gb[f] : ga[f]);
  }
  if (rc == LBM_OK && two) {
    // the previous level through the second copy, then into sb3 / psb
    if ((rc = ensure_aux(p)) == LBM_OK) {
      const bool have_raw = p->have_raw, raw_pending = p->raw_pending;
      swap_state(p);
      rc = set_state_t<T>(p, b[0], b[1], b[2], b[3], b[4], st);
      swap_state(p);
      p->have_raw = have_raw;
      p->raw_pending = raw_pending;
    }
  }
  if (rc == LBM_OK) rc = set_state_t<T>(p, a[0], a[1], a[2], a[3], a[4], st);
  if (rc == LBM_OK && two) {
    const size_t cs = 2 * p->rsize(), rawbytes = p->rsize() * 4 * n3;
    if (!p->rawb && hipMalloc(&p->rawb, rawbytes) != hipSuccess) rc = fail(LBM_ERR_RUNTIME, "allocation failed");
    if (rc == LBM_OK &&
        (hipMemcpyAsync(p->rawb, p->x_raw, rawbytes, hipMemcpyDeviceToDevice, st) != hipSuccess ||
         hipMemcpyAsync(p->sb3, p->x_sa3, cs * p->nspec * 4 * E * L, hipMemcpyDeviceToDevice, st) != hipSuccess ||
        hipMemcpyAsync(p->psb, p->x_psa, p->rsize() * n2, hipMemcpyDeviceToDevice, st) != hipSuccess ||
         (p->sb2 && hipMemcpyAsync(p->sb2, p->x_sa2, cs * p->nspec * 3 * E, hipMemcpyDeviceToDevice, st) != hipSuccess)))
      rc = fail(LBM_ERR_RUNTIME, "previous-level copy failed");
  }
  if (hipStreamSynchronize(st) != hipSuccess && rc == LBM_OK) rc = fail(LBM_ERR_RUNTIME, "sync failed");
  (void)hipFree(buf);
  return rc;
}

// ---------------------------------------------------------------------------
// Spectral initial state, linear response and the adjoint sweep.
// ---------------------------------------------------------------------------
template <class T>
int set_spectral_t(lbm_plan *p, const double *vor, const double *div, const double *t,
                   const double *q, const double *pi, hipStream_t st) {
  const int J = p->J, I = p->I, L = p->L, E = p->E, M = p->M;
  const int EL = E * L;
  const size_t s3 = (size_t)p->nspec * 4 * EL, s2 = (size_t)p->nspec * 3 * E;
  HIPCHK(hipMemsetAsync(p->pend, 0, E * sizeof(int), st));
  const double *src[4] = {vor, div, t, q};
  HIPCHK(hipMemsetAsync(p->sa3, 0, s3 * sizeof(cx<T>), st));
  HIPCHK(hipMemsetAsync(p->sa2, 0, s2 * sizeof(cx<T>), st));
  const size_t n3 = (size_t)p->nspec * EL;
  for (int f = 0; f < 4; ++f)
    if (src[f])
      hipLaunchKernelGGL((k_pack_spec<T>), dim3(nblk(n3)), dim3(256), 0, st, src[f], (cx<T> *)p->sa3,
                         E, M, L, 4, f, EL);
  if (pi)
    hipLaunchKernelGGL((k_pack_spec<T>), dim3(nblk((size_t)p->nspec * E)), dim3(256), 0, st, pi,
                       (cx<T> *)p->sa2, E, M, 1, 3, 0, E);
  // no global mean of vorticity or divergence (the reference's grid state
  // holds winds: state.py:94-99); the tangent-linear model keeps its input
  if (!p->linear) HIPCHK(hipMemsetAsync(p->sa3, 0, 2 * (size_t)EL * sizeof(cx<T>), st));
  hipLaunchKernelGGL((k_vsum<T>), dim3(nblk((size_t)p->nspec * E)), dim3(256), 0, st,
                     (const cx<T> *)p->sa3, (cx<T> *)p->sa2, p->tabs<T>().vc, E, L, p->nspec);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(p->sb3, p->sa3, s3 * sizeof(cx<T>), hipMemcpyDeviceToDevice, st));
  if (p->sb2) HIPCHK(hipMemcpyAsync(p->sb2, p->sa2, s2 * sizeof(cx<T>), hipMemcpyDeviceToDevice, st));
  const size_t n2 = (size_t)E * J * I;
  if (p->linear || !pi) {
    HIPCHK(hipMemsetAsync(p->psa, 0, n2 * sizeof(T), st));
    HIPCHK(hipMemsetAsync(p->psb, 0, n2 * sizeof(T), st));
  } else {
    // ps' of both levels from the synthesised pi (state.py:94-99)
    const ColMap cm{E, 1, 0, 1};
    const size_t gb = (n2 * sizeof(T) + 255) / 256 * 256;
    void *base;
    CHK(scratch_get(p, gb + (size_t)M * J * fpad(E) * sizeof(cx<T>), &base, st));
    T *g = (T *)base;
    cx<T> *F = (cx<T> *)((char *)base + gb);
    OutSlots sl{{0}};
    CHK((launch_syn<T, SynScalar>(p, p->sa2, F, E, E, sl, st, 3)));
    CHK(launch_ifft<T>(p, F, g, cm, st));
    const BRow br{L};
    hipLaunchKernelGGL((k_ps_from_pi<T>), dim3(nblk(n2)), dim3(256), 0, st, (const T *)g, 1, 0,
                       (const T *)p->bsg, br.rows(), br.psbar(), (T *)p->psa, (T *)p->psb, E, J, I,
                       p->nb);
    HIPCHK(hipGetLastError());
  }
  p->have_raw = false;
  p->raw_pending = false;
  return LBM_OK;
}

template <class T>
int set_response_t(lbm_plan *p, const double *wgrid, const double *wlev, const double *wstep,
                   int nst, hipStream_t st) {
  const int J = p->J, I = p->I, L = p->L, M = p->M;
  if (!p->obs) HIPCHK(hipMalloc(&p->obs, sizeof(cx<T>) * p->nspec));
  if (!p->wlev) HIPCHK(hipMalloc(&p->wlev, sizeof(T) * L));
  if (!p->jacc) HIPCHK(hipMalloc(&p->jacc, sizeof(double) * p->E));
  if (nst > p->wstep_cap) {
    if (p->wstep) {
      HIPCHK(hipStreamSynchronize(st));
      HIPCHK(hipFree(p->wstep));
      p->wstep = nullptr;
    }
    HIPCHK(hipMalloc(&p->wstep, sizeof(T) * std::max(nst, 1)));
    p->wstep_cap = nst;
  }
  std::vector<T> wl(wlev, wlev + L), ws(wstep, wstep + nst);
  HIPCHK(hipMemcpy(p->wlev, wl.data(), sizeof(T) * L, hipMemcpyHostToDevice));
  if (nst > 0) HIPCHK(hipMemcpy(p->wstep, ws.data(), sizeof(T) * nst, hipMemcpyHostToDevice));
  p->nwstep = nst;
  // obs = (wave_to_grid o inverse Laplacian)^T of the grid weights
  const size_t ng = (size_t)J * I;
  const size_t al = 256;
  auto up = [&](size_t b) { return (b + al - 1) / al * al; };
  void *base;
  CHK(scratch_get(p, up(ng * sizeof(double)) + up(ng * sizeof(T)) + (size_t)M * J * fpad(1) * sizeof(cx<T>),
                  &base, st));
  double *w64 = (double *)base;
  T *g = (T *)((char *)base + up(ng * sizeof(double)));
  cx<T> *F = (cx<T> *)((char *)base + up(ng * sizeof(double)) + up(ng * sizeof(T)));
  HIPCHK(hipMemcpyAsync(w64, wgrid, ng * sizeof(double), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL((k_pack_grid<T>), dim3(nblk(ng)), dim3(256), 0, st, (const double *)w64, g, 1,
                     J, I, 1, 1, 0, (const double *)nullptr);
  HIPCHK(hipGetLastError());
  const ColMap cm{1, 1, 1, 0};
  CHK(launch_fft<T>(p, g, F, cm, st, 2));
  DevTabs<T> &t = p->tabs<T>();
  InSlots sl{{0}};
  CHK((launch_ana<T, AdjPsi>(p, F, p->obs, 1, 1, sl, st, 1, 0, t.Ps, t.Hs, t.inv_lap)));
  HIPCHK(hipMemsetAsync(p->jacc, 0, sizeof(double) * p->E, st));
  HIPCHK(hipStreamSynchronize(st));
  p->resp_set = true;
  return LBM_OK;
}

// One adjoint step: transposed spectral step, transposed analysis (synthesis
// kernels with the analysis tables), inverse FFT, transposed tendencies,
// forward FFT, transposed synthesis (analysis kernels with the synthesis
// tables), then the step counter back by one.
template <class T>
int launch_step_adj(lbm_plan *p, hipStream_t st, int ecnt) {
  DevTabs<T> &t = p->tabs<T>();
  StepAdjArgs<T> a;
  a.aa3 = (cx<T> *)p->aa3;
  a.ab3 = (cx<T> *)p->ab3;
  a.aa2 = (cx<T> *)p->aa2;
  a.ab2 = (cx<T> *)p->ab2;
  a.s3a = (const cx<T> *)p->s3a;
  a.at3 = (cx<T> *)p->tend3;
  a.at2 = (cx<T> *)p->tend2;
  a.kv = t.kv;
  a.kd = t.kd;
  a.ktt = t.ktt;
  a.kq = t.kq;
  a.minv = t.minv;
  a.vinv = t.vinv;
  a.vc = t.vc;
  a.tref = t.tref;
  a.lap = t.lap;
  a.n_of_s = p->d_n_of_s;
  a.obs = p->resp_set ? (const cx<T> *)p->obs : nullptr;
  a.wlev = (const T *)p->wlev;
  a.wstep = (const T *)p->wstep;
  a.nwstep = p->nwstep;
  a.ctl = p->ctl;
  a.nsteps = p->d_nsteps;
  a.E = p->E;
  a.M = p->M;
  a.nspec = p->nspec;
  a.nbasic = p->nb;
  a.inistp = p->inistp;
  a.vmask = p->vmask;
  a.skip_low_n = p->skip_low_n;
  a.delt = (T)p->delt;
  a.robert = (T)p->robert;
  a.L = p->L;
  a.ecnt = ecnt;
  a.traj = p->traj_cur != nullptr;
  const size_t nthr = (size_t)p->nspec * ecnt;
  // two roles per (s, member): see k_step_adj; at 20 levels the lane
  // mappings of the forward step (one total wavenumber per wave) except
  // where every wave holds one coefficient and one basic state already
  const StepMap map = p->L == 20 ? step_map(p, ecnt) : STEP_LANES;
  const bool ng = map == STEP_NG || map == STEP_NGM;
  const dim3 grid(ng ? step_ng_waves(p->M, ecnt) : (unsigned)((nthr + 63) / 64), 2);
  if (p->L == 20 && map == STEP_NG)
    hipLaunchKernelGGL((k_step_adj<T, 20, true, 1>), grid, dim3(64), 0, st, a);
  else if (p->L == 20 && map == STEP_NGM)
    hipLaunchKernelGGL((k_step_adj<T, 20, true, 128 / (2 * sizeof(T))>), grid, dim3(64), 0, st, a);
  else if (p->L == 20)
    hipLaunchKernelGGL((k_step_adj<T, 20>), grid, dim3(64), 0, st, a);
  else if (p->L == 40)
    hipLaunchKernelGGL((k_step_adj<T, 40>), grid, dim3(64), 0, st, a);
  else
    HIPCHK(launch_step_adj_wide<T>(a, grid, st));
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

template <class T>
int adjoint_step(lbm_plan *p, hipStream_t st, int ecnt) {
  const int E = p->E, L = p->L, EL = E * L;
  DevTabs<T> &t = p->tabs<T>();
  const ActiveMask am = loop_mask(p, false, ecnt);
  const int ne = ecnt > 0 ? ecnt : E;
  // The surface synthesis and analysis run on the second stream beside the
  // level ones (disjoint outputs), joined before their first consumers (the
  // inverse FFT, the end of the step), as in the forward step: launches of
  // more than 8 members, never while profiling.
  const bool fork = !p->prof && p->side != nullptr && ne > 8;
  hipStream_t sd = fork ? p->side : st;
  auto fork_at = [&](int k) -> int {
    if (fork) {
      HIPCHK(hipEventRecord(p->fj[k], st));
      HIPCHK(hipStreamWaitEvent(sd, p->fj[k], 0));
    }
    return LBM_OK;
  };
  auto join_side = [&](int k) -> int {
    if (fork) {
      HIPCHK(hipEventRecord(p->fj[k], sd));
      HIPCHK(hipStreamWaitEvent(st, p->fj[k], 0));
    }
    return LBM_OK;
  };
  CHK(mark(p, 7, st));
  CHK(launch_step_adj<T>(p, st, ne));
  const ColMap cmo{E, L, 9, 2}, cmi{E, L, 6, 6};
  const int Co = cmo.ncols(), Ci = cmi.ncols();
  CHK(fork_at(0));
  CHK(mark(p, 5, st));
  {
    OutSlots sl{{0, EL, 2 * EL, 3 * EL, 4 * EL, 5 * EL, 6 * EL, 7 * EL, 8 * EL}};
    CHK((launch_syn<T, AdjTend>(p, p->tend3, p->fout, EL, Co, sl, st, 4, t.Pa, t.Ha, t.nn1a2, am)));
  }
  CHK(mark(p, 6, st));
  {
    // the default model also carries pi_prev* (tendency field 1): both
    // surface fields as 2E columns, as the forward analysis runs them
    OutSlots sl{{9 * EL}};
    const bool traj = p->traj_cur != nullptr;
    CHK((launch_syn<T, AdjScalarFft>(p, p->tend2, p->fout, traj ? 2 * E : E, Co, sl, sd,
                                     traj ? 1 : 2, t.Pa, t.Ha, t.nn1a2, am)));
  }
  CHK(join_side(1));
  CHK(mark(p, 4, st));
  CHK(launch_ifft<T>(p, p->fout, p->gout, cmo, st, 1, am));
  CHK(mark(p, 3, st));
  TendAdjArgs<T> ta;
  ta.gout_a = (const T *)p->gout;
  ta.gin_a = (T *)p->gin;
  ta.frc_a = (T *)p->frc_a;
  ta.bsg = (const T *)p->bsg;
  ta.vc = t.vc;
  ta.tref = t.tref;
  ta.rcos = t.rcos;
  ta.E = E;
  ta.J = p->J;
  ta.I = p->I;
  ta.L = L;
  ta.nbasic = p->nb;
  ta.am = am;
  ta.traj = (const T *)p->traj_cur;
  ta.psa_a = (T *)p->psa_a;
  ta.psb_a = (T *)p->psb_a;
  ta.nsteps = p->d_nsteps;
  ta.inistp = p->inistp;
  ta.step = p->traj_step;
  ta.full_ke = !(p->strictly_linear || p->linear);
  ta.robert = (T)p->robert;
  hipLaunchKernelGGL((k_tend_adj<T>), dim3(ne * p->J * (p->I / grid_chunk(p->I))),
                     dim3(grid_chunk(p->I)), 0, st, ta);
  HIPCHK(hipGetLastError());
  CHK(mark(p, 2, st));
  CHK(launch_fft<T>(p, p->gin, p->fin, cmi, st, 2, am));
  CHK(fork_at(2));
  CHK(mark(p, 0, st));
  {
    InSlots sl{{0, EL, 2 * EL, 3 * EL, 4 * EL, 5 * EL}};
    CHK((launch_ana<T, AdjSynLevel>(p, p->fin, p->aa3, EL, Ci, sl, st, 4, 0, t.Ps, t.Hs, t.inv_lap,
                                    am)));
  }
  CHK(mark(p, 1, st));
  {
    InSlots sl{{6 * EL, 6 * EL + E, 6 * EL + 2 * E, 6 * EL + 3 * E, 6 * EL + 4 * E, 6 * EL + 5 * E}};
    CHK((launch_ana<T, AdjSynSurf>(p, p->fin, p->s3a, E, Ci, sl, sd, 3, 0, t.Ps, t.Hs, t.inv_lap,
                                   am)));
  }
  CHK(join_side(3));
  CHK(mark(p, 8, st));
  hipLaunchKernelGGL(k_tick_back, dim3(1), dim3(1), 0, st, p->ctl);
  HIPCHK(hipGetLastError());
  if (p->prof) {
    // slots keep the forward meaning of the transposed kernel
    const int order[8] = {7, 5, 6, 4, 3, 2, 0, 1};
    const int next[8] = {5, 6, 4, 3, 2, 0, 1, 8};
    HIPCHK(hipEventSynchronize(p->ev[8]));
    for (int k = 0; k < 8; ++k) {
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, p->ev[order[k]], p->ev[next[k]]));
      p->prof_ms[order[k]] += ms;
      p->prof_n[order[k]] += 1;
    }
  }
  return LBM_OK;
}

int adjoint_dispatch(lbm_plan *p, hipStream_t st, int ecnt) {
  return p->prec == LBM_FP64 ? adjoint_step<double>(p, st, ecnt) : adjoint_step<float>(p, st, ecnt);
}

template <class T>
int adjoint_finish(lbm_plan *p, double *gvor, double *gdiv, double *gt, double *gq, double *gpi,
                   double *gfrc, hipStream_t st) {
  const int E = p->E, L = p->L, M = p->M, J = p->J, I = p->I;
  const size_t EL = (size_t)E * L;
  hipLaunchKernelGGL((k_vsum_adj<T>), dim3(nblk((size_t)p->nspec * E)), dim3(256), 0, st,
                     (const cx<T> *)p->s3a, (cx<T> *)p->aa3, (cx<T> *)p->aa2, p->tabs<T>().vc, E, L,
                     p->nspec);
  // the initial state fills both time levels: gradient = Sa* + Sb*
  const size_t n3 = (size_t)p->nspec * 4 * EL, n2 = (size_t)p->nspec * E;
  hipLaunchKernelGGL((k_add_cx<T>), dim3(nblk(n3)), dim3(256), 0, st, (const cx<T> *)p->aa3,
                     (const cx<T> *)p->ab3, (cx<T> *)p->aa3, n3);
  hipLaunchKernelGGL((k_add_cx<T>), dim3(nblk(n2)), dim3(256), 0, st, (const cx<T> *)p->aa2,
                     (const cx<T> *)p->ab2, (cx<T> *)p->aa2, n2);
  // the default model drops the global means of the initial vorticity and
  // divergence on entry (coefficient s = 0 of fields 0, 1): no gradient
  if (!p->linear) HIPCHK(hipMemsetAsync(p->aa3, 0, 2 * EL * sizeof(cx<T>), st));
  double *dst[4] = {gvor, gdiv, gt, gq};
  const size_t tot = (size_t)E * M * M * L;
  for (int f = 0; f < 4; ++f)
    if (dst[f])
      hipLaunchKernelGGL((k_unpack_spec<T>), dim3(nblk(tot)), dim3(256), 0, st,
                         (const cx<T> *)p->aa3, dst[f], E, M, L, 4, f, (int)EL);
  if (gpi)
    hipLaunchKernelGGL((k_unpack_spec<T>), dim3(nblk((size_t)E * M * M)), dim3(256), 0, st,
                       (const cx<T> *)p->aa2, gpi, E, M, 1, 1, 0, E);
  if (gfrc)
    hipLaunchKernelGGL((k_unpack_grid<T>), dim3(nblk((size_t)E * J * I * L)), dim3(256), 0, st,
                       (const T *)p->frc_a, gfrc, E, J, I, L, L, 0, 0, (const T *)nullptr, 0, 0, 1);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

int capture_adjoint(lbm_plan *p, int kind, int nsteps, int ecnt, hipStream_t st,
                    hipGraphExec_t *out) {
  if (p->prof) return fail(LBM_ERR_STATE, "profiling runs eager launches only (use_graph=0)");
  const auto key = std::make_pair(kind, ecnt);
  auto it = p->graphs.find(key);
  if (it != p->graphs.end()) {
    *out = it->second;
    return LBM_OK;
  }
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int rc = LBM_OK;
  for (int s = 0; s < nsteps && rc == LBM_OK; ++s) rc = adjoint_dispatch(p, st, ecnt);
  hipError_t e2 = hipStreamEndCapture(st, &g);
  if (rc != LBM_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e2 != hipSuccess) return fail(LBM_ERR_RUNTIME, "graph capture failed: %s", hipGetErrorString(e2));
  hipGraphExec_t ge = nullptr;
  hipError_t e3 = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e3 != hipSuccess) return fail(LBM_ERR_RUNTIME, "graph instantiate failed: %s", hipGetErrorString(e3));
  p->graphs[key] = ge;
  *out = ge;
  return LBM_OK;
}

template <class T>
size_t alloc_bytes(const lbm_plan *p) {
  (void)p;
  return sizeof(T);
}

// ---------------------------------------------------------------------------
// History extraction: synthesis of the spectral sums, unpacking to the
// reference's grid layout.
// ---------------------------------------------------------------------------
// (vor, div, t, z) spectral [s][4][L][E] in p->tend3 -> grid rows
// (psi, chi, u, v, t, z) [e][j][6L][I] in p->gin -> the outputs that are
// not null, (E, J, I, L) float64.
template <class T>
int hist_synth(lbm_plan *p, double *const out[6], hipStream_t st) {
  const int E = p->E, L = p->L, J = p->J, I = p->I, EL = E * L;
  const ColMap cm{E, L, 6, 0};
  OutSlots sl{{0, EL, 2 * EL, 3 * EL, 4 * EL, 5 * EL}};
  CHK((launch_syn<T, SynHist>(p, p->tend3, p->fin, EL, cm.ncols(), sl, st, SynHist::NIN)));
  CHK(launch_ifft<T>(p, p->fin, p->gin, cm, st));
  const size_t n3 = (size_t)E * J * I * L;
  for (int f = 0; f < 6; ++f)
    if (out[f])
      hipLaunchKernelGGL((k_unpack_grid<T>), dim3(nblk(n3)), dim3(256), 0, st, (const T *)p->gin,
                         out[f], E, J, I, L, cm.rows(), f * L, 0, (const T *)nullptr, 0, 0, 1);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

template <class T>
int get_history_t(lbm_plan *p, int slot, double *psi, double *chi, double *u, double *v, double *w,
                  double *t, double *z, double *ps, hipStream_t st) {
  const int E = p->E, L = p->L, J = p->J, I = p->I;
  const T scale = p->hist_mode == 1 ? (T)1 / (T)p->hist_K : (T)1;
  hipLaunchKernelGGL((k_hist_spec_out<T>), dim3(nblk((size_t)p->nspec * E)), dim3(256), 0, st,
                     (cx<T> *)p->hspec[slot], (cx<T> *)p->tend3, p->tabs<T>().vc, E, L, p->nspec,
                     scale, 1);
  HIPCHK(hipGetLastError());
  double *const out[6] = {psi, chi, u, v, t, z};
  CHK(hist_synth<T>(p, out, st));
  hipLaunchKernelGGL((k_hist_grid_out<T>), dim3(nblk((size_t)E * J * (L + 1) * I)), dim3(256), 0, st,
                     (T *)p->hgrid[slot], w, ps, E, J, I, L, scale, 1);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

template <class T>
int get_diagnostics_t(lbm_plan *p, double *psi, double *chi, double *w, double *z, hipStream_t st) {
  const int E = p->E, L = p->L, J = p->J, I = p->I;
  hipLaunchKernelGGL((k_diag_spec<T>), dim3(nblk((size_t)p->nspec * E)), dim3(256), 0, st,
                     (const cx<T> *)p->hlast_spec, (const cx<T> *)p->sa3, (cx<T> *)p->tend3,
                     p->tabs<T>().vc, E, L, p->nspec);
  HIPCHK(hipGetLastError());
  double *const out[6] = {psi, chi, nullptr, nullptr, nullptr, z};
  CHK(hist_synth<T>(p, out, st));
  if (w)
    hipLaunchKernelGGL((k_unpack_grid<T>), dim3(nblk((size_t)E * J * I * L)), dim3(256), 0, st,
                       (const T *)p->hlast_grid, w, E, J, I, L, L, 0, 0, (const T *)nullptr, 0, 0, 1);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

// Analyses of the steady forcings (nf members), packed for k_add_steady.
template <class T>
int set_steady_t(lbm_plan *p, const double *vor, const double *div, const double *ps, int nf,
                 hipStream_t st) {
  const int M = p->M, L = p->L;
  const size_t nv = (size_t)p->nspec * 2 * L * nf, np = (size_t)p->nspec * nf;
  double *w = nullptr;
  HIPCHK(hipMalloc(&w, sizeof(double) * 2 * nf * M * M * L));
  int rc = LBM_OK;
  if (vor || div) {
    rc = hipMalloc(&p->sfv, nv * sizeof(cx<T>)) == hipSuccess ? LBM_OK : fail(LBM_ERR_RUNTIME, "out of memory");
    if (rc == LBM_OK && hipMemsetAsync(p->sfv, 0, nv * sizeof(cx<T>), st) != hipSuccess)
      rc = fail(LBM_ERR_RUNTIME, "initialisation failed");
    const double *src[2] = {vor, div};
    for (int f = 0; f < 2 && rc == LBM_OK; ++f) {
      if (!src[f]) continue;
      rc = spectral_op<T>(p, OP_G2W, src[f], nullptr, w, nullptr, nf, L, st);
      if (rc == LBM_OK)
        hipLaunchKernelGGL((k_pack_spec<T>), dim3(nblk((size_t)p->nspec * nf * L)), dim3(256), 0,
                           st, w, (cx<T> *)p->sfv, nf, M, L, 2, f, nf * L);
    }
  }
  if (ps && rc == LBM_OK) {
    rc = hipMalloc(&p->sfp, np * sizeof(cx<T>)) == hipSuccess ? LBM_OK : fail(LBM_ERR_RUNTIME, "out of memory");
    if (rc == LBM_OK) rc = spectral_op<T>(p, OP_G2W, ps, nullptr, w, nullptr, nf, 1, st);
    if (rc == LBM_OK)
      hipLaunchKernelGGL((k_pack_spec<T>), dim3(nblk(np)), dim3(256), 0, st, w, (cx<T> *)p->sfp,
                         nf, M, 1, 1, 0, nf);
  }
  if (hipGetLastError() != hipSuccess && rc == LBM_OK) rc = fail(LBM_ERR_RUNTIME, "kernel launch failed");
  (void)hipStreamSynchronize(st);
  (void)hipFree(w);
  return rc;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
namespace {

// One copy of the whole model state (both spectral levels, pi and the
// column sums, ps' of both levels, the pending ps' update, loop control).
struct StateParts {
  size_t off[7], size[7], total;
};
StateParts state_parts(const lbm_plan *p) {
  const size_t cs = 2 * p->rsize(), EL = (size_t)p->E * p->L;
  const size_t sz[7] = {cs * p->nspec * 4 * EL, cs * p->nspec * 4 * EL, cs * p->nspec * 3 * (size_t)p->E,
                        p->rsize() * (size_t)p->E * p->J * p->I, p->rsize() * (size_t)p->E * p->J * p->I,
                        sizeof(int) * (size_t)p->E, sizeof(Ctl)};
  StateParts sp;
  size_t o = 0;
  for (int k = 0; k < 7; ++k) {
    sp.off[k] = o;
    sp.size[k] = sz[k];
    o += (sz[k] + 255) / 256 * 256;
  }
  sp.total = o;
  return sp;
}
size_t state_bytes(const lbm_plan *p) { return state_parts(p).total; }

int copy_state(lbm_plan *p, void *buf, bool save, hipStream_t st) {
  const StateParts sp = state_parts(p);
  void *parts[7] = {p->sa3, p->sb3, p->sa2, p->psa, p->psb, p->pend, p->ctl};
  for (int k = 0; k < 7; ++k) {
    char *b = (char *)buf + sp.off[k];
    if (save)
      HIPCHK(hipMemcpyAsync(b, parts[k], sp.size[k], hipMemcpyDeviceToDevice, st));
    else
      HIPCHK(hipMemcpyAsync(parts[k], b, sp.size[k], hipMemcpyDeviceToDevice, st));
  }
  return LBM_OK;
}

int grow(void **buf, size_t *cap, size_t bytes, hipStream_t st) {
  if (bytes <= *cap) return LBM_OK;
  if (*buf) {
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipFree(*buf));
    *buf = nullptr;
    *cap = 0;
  }
  HIPCHK(hipMalloc(buf, bytes));
  *cap = bytes;
  return LBM_OK;
}

}  // namespace

namespace {

// The grid field (ps_bar + ps'_0) (ps'_a* + ps'_b*) of the initial state:
// ps' of both levels was ps_bar expm1(pi_0) on the grid, so this is the
// grid adjoint of pi_0 whose synthesis transpose joins the pi gradient.
template <class T>
__global__ void k_ps0_adj(const T *__restrict__ traj0, const T *__restrict__ bsg,
                          const T *__restrict__ psa_a, const T *__restrict__ psb_a,
                          double *__restrict__ out, int E, int J, int I, int L, int nbasic) {
  const size_t tot = (size_t)E * J * I;
  const BRow br{L};
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot;
       q += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % I);
    const size_t ej = q / I;
    const int j = (int)(ej % J), e = (int)(ej / J);
    const int eb = nbasic > 1 ? e : 0;
    const T pb = bsg[(((size_t)eb * J + j) * br.rows() + br.psbar()) * I + i];
    const T ps0 = traj0[(ej * (2 * L + 2) + 2 * L) * I + i];
    out[q] = (double)((pb + ps0) * (psa_a[q] + psb_a[q]));
  }
}

__global__ void k_add_f64(const double *__restrict__ x, double *__restrict__ y, size_t n) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x)
    y[q] += x[q];
}

// Adjoint of the default model (not tangent-linear): backward over the
// checkpointed segments of the last run; each segment is recomputed from
// its checkpoint storing the per-step trajectory, then swept backward with
// the transposed kernels about that trajectory.
template <class T>
int adjoint_traj_t(lbm_plan *p, const std::vector<int> &len, double *const out[6],
                   hipStream_t st) {
  const int E = p->E, L = p->L, J = p->J, I = p->I, K = p->ckpt_every;
  const int total = p->ckpt_total;
  const size_t slot = p->rsize() * (size_t)E * J * (2 * L + 2) * I;
  const size_t nps = p->rsize() * (size_t)E * J * I;
  CHK(grow(&p->traj, &p->traj_cap, (size_t)K * slot, st));
  size_t keep_cap = p->keep ? state_parts(p).total : 0;
  CHK(grow(&p->keep, &keep_cap, state_parts(p).total, st));
  for (void **b : {&p->psa_a, &p->psb_a}) {
    if (!*b) HIPCHK(hipMalloc(b, nps));
    HIPCHK(hipMemsetAsync(*b, 0, nps, st));
  }
  CHK(copy_state(p, p->keep, true, st));
  // the recomputed steps add no response, history or diagnostics
  const bool resp = p->resp_set, keep_last = p->keep_last;
  const int hist_mode = p->hist_mode;
  p->resp_set = false;
  p->hist_mode = 0;
  p->keep_last = false;
  int rc = LBM_OK;
  for (int g = p->ckpt_count - 1; g >= 0 && rc == LBM_OK; --g) {
    const int k0 = g * K, k1 = std::min(total, k0 + K);
    rc = copy_state(p, (char *)p->ckpt + (size_t)g * state_parts(p).total, false, st);
    for (int k = k0; k < k1 && rc == LBM_OK; ++k) {
      p->traj_cur = (char *)p->traj + (size_t)(k - k0) * slot;
      rc = step_dispatch(p, st, all_live(p, len, k) ? 0 : live_members(p, len, k, true));
    }
    if (rc != LBM_OK) break;
    p->resp_set = resp;
    static Ctl cb;  // host source of the copy below, alive until it is done
    cb = Ctl{k1 - 1, 0, INT_MAX, 0};
    if (hipMemcpyAsync(p->ctl, &cb, sizeof(Ctl), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      rc = fail(LBM_ERR_RUNTIME, "control block upload failed");
      break;
    }
    for (int k = k1 - 1; k >= k0 && rc == LBM_OK; --k) {
      p->traj_cur = (char *)p->traj + (size_t)(k - k0) * slot;
      p->traj_step = k;
      rc = adjoint_dispatch(p, st, all_live(p, len, k) ? 0 : live_members(p, len, k, false));
    }
    p->resp_set = false;
  }
  p->traj_cur = nullptr;
  p->traj_step = -1;
  p->resp_set = resp;
  p->hist_mode = hist_mode;
  p->keep_last = keep_last;
  if (rc == LBM_OK)
    rc = adjoint_finish<T>(p, out[0], out[1], out[2], out[3], out[4], out[5], st);
  if (rc == LBM_OK && out[4]) {
    // pi_0 also set ps' of both levels on the grid (segment 0's trajectory
    // slot 0 still holds the initial ps')
    const size_t n2 = (size_t)E * J * I, nw = (size_t)E * p->M * p->M * 2;
    double *g = nullptr, *w = nullptr;
    if (hipMalloc(&g, n2 * sizeof(double)) != hipSuccess || hipMalloc(&w, nw * sizeof(double)) != hipSuccess)
      rc = fail(LBM_ERR_RUNTIME, "allocation failed");
    if (rc == LBM_OK) {
      hipLaunchKernelGGL((k_ps0_adj<T>), dim3(nblk(n2)), dim3(256), 0, st, (const T *)p->traj,
                         (const T *)p->bsg, (const T *)p->psa_a, (const T *)p->psb_a, g, E, J, I, L,
                         p->nb);
      rc = spectral_op<T>(p, OP_W2G_ADJ, g, nullptr, w, nullptr, E, 1, st);
    }
    if (rc == LBM_OK) {
      hipLaunchKernelGGL(k_add_f64, dim3(nblk(nw)), dim3(256), 0, st, (const double *)w, out[4], nw);
      if (hipGetLastError() != hipSuccess) rc = fail(LBM_ERR_RUNTIME, "launch failed");
    }
    if (hipStreamSynchronize(st) != hipSuccess && rc == LBM_OK) rc = fail(LBM_ERR_RUNTIME, "sync failed");
    if (g) (void)hipFree(g);
    if (w) (void)hipFree(w);
  }
  // the forward state is left as the run left it
  if (copy_state(p, p->keep, false, st) != LBM_OK && rc == LBM_OK) rc = LBM_ERR_RUNTIME;
  return rc;
}

}  // namespace

// ---------------------------------------------------------------------------
// Device scratch of lbm_grid_stats_noalloc, one copy per GPU for the life of
// the process: 256 bytes for the result pair, then kStatsSlots block pairs
// (64 KiB + 256 bytes).  Host code reaches it only through the address
// hipGetSymbolAddress returns on each device.
// ---------------------------------------------------------------------------
namespace {
constexpr int kStatsSlots = 4096;  // block (max, count) pairs
constexpr int kStatsOff = 32;      // the pairs start 256 bytes in
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned long long scratch[kStatsOff + 2 * kStatsSlots];

namespace {
std::mutex g_scratch_mu;
unsigned long long *g_scratch[kMaxDevices];  // scratch's address on each device

// The current device's copy of scratch, looked up on the first call there.
int device_scratch(unsigned long long **out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(LBM_ERR_RUNTIME, "device %d is beyond the %d the library tracks", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(LBM_ERR_RUNTIME, "no device address for the statistics scratch");
    g_scratch[dev] = (unsigned long long *)ptr;
  }
  *out = g_scratch[dev];
  return LBM_OK;
}
}  // namespace

extern "C" {

int lbm_abi_version(void) { return LBM_ABI_VERSION; }

int lbm_last_error(char *buf, int64_t len) {
  if (!buf || len <= 0) return LBM_ERR_VALUE;
  std::snprintf(buf, (size_t)len, "%s", g_err.c_str());
  return LBM_OK;
}

int lbm_plan_destroy(lbm_plan *p) {
  if (!p) return LBM_OK;
  // nothing of the plan may still run when its graphs, buffers and streams go
  if (p->own) (void)hipStreamSynchronize(p->own);
  if (p->side) (void)hipStreamSynchronize(p->side);
  destroy_graphs(p);
  free_tabs(p->d64);
  free_tabs(p->d32);
  void *bufs[] = {p->d_n_of_s, p->d_mu, p->d_psbar, p->sa3, p->sa2, p->sb3, p->tend3, p->tend2,
                  p->fin, p->gin, p->gout, p->fout, p->raw, p->psa, p->psb, p->bsg, p->frc,
                  p->ctl, p->pend, p->d_nsteps, p->scratch, p->sb2, p->aa3, p->ab3, p->aa2,
                  p->ab2, p->s3a, p->frc_a, p->obs, p->wlev, p->wstep, p->jacc,
                  p->hspec[0], p->hspec[1], p->hgrid[0], p->hgrid[1], p->hlast_spec,
                  p->hlast_grid, p->mjo, p->sfv, p->sfp, p->x_sa3, p->x_sa2, p->x_sb3,
                  p->x_sb2, p->x_psa, p->x_psb, p->x_raw, p->x_pend, p->x_ctl, p->rawb, p->ckpt, p->traj, p->keep, p->psa_a,
                  p->psb_a};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  for (auto &e : p->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : p->fj)
    if (e) (void)hipEventDestroy(e);
  if (p->side) (void)hipStreamDestroy(p->side);
  if (p->own) (void)hipStreamDestroy(p->own);
  delete p;
  return LBM_OK;
}

int lbm_plan_create(const lbm_desc *d, lbm_plan **out) {
  if (!d || !out) return fail(LBM_ERR_VALUE, "null descriptor or output");
  *out = nullptr;
  FftFactors ffac{};
  if (d->nlon != 128 && d->nlon != 384) {
    // the Stockham path: radices 4, 2, 3, 5 of nlon / 2, at most 512
    bool ok = d->nlon >= 64 && d->nlon % 64 == 0 && d->nlon / 2 <= 512;
    int n = ok ? d->nlon / 2 : 1;
    for (int f : {4, 2, 3, 5})
      while (n % f == 0 && ffac.np < 12) {
        ffac.r[ffac.np++] = f;
        n /= f;
      }
    if (!ok || n != 1)
      return fail(LBM_ERR_VALUE, "nlon=%d: longitudes must be a multiple of 64, at most 1024, "
                  "with nlon/2 a product of 2, 3 and 5", d->nlon);
    ffac.NI = d->nlon;
    ffac.NC = d->nlon / 2;
    for (int q = 0, ns = 1; q < ffac.np; ns *= ffac.r[q], ++q) {
      ffac.ns[q] = ns;
      ffac.dmul[q] = ((1 << 20) + ns - 1) / ns;
    }
  }
  if (d->nlat <= 0 || d->nlat % 8 || d->nlat > 64 * 32)
    return fail(LBM_ERR_VALUE, "nlat=%d: a positive multiple of 8 is required", d->nlat);
  if (d->nmax < 1 || d->nmax + 1 > d->nlon / 2 || d->nmax >= d->nlat)
    return fail(LBM_ERR_VALUE, "nmax=%d unsupported for nlon=%d, nlat=%d", d->nmax, d->nlon,
                d->nlat);
  if (d->nlev < 1 || d->nlev > LBM_MAX_LEVELS)
    return fail(LBM_ERR_VALUE, "nlev=%d: 1 to %d levels are supported", d->nlev, LBM_MAX_LEVELS);
  if (d->nmem < 1) return fail(LBM_ERR_VALUE, "nmem must be >= 1");
  if (d->nbasic != 1 && d->nbasic != d->nmem) return fail(LBM_ERR_VALUE, "nbasic must be 1 or nmem");
  if (d->precision != LBM_FP64 && d->precision != LBM_FP32 && d->precision != LBM_BF16X3)
    return fail(LBM_ERR_VALUE, "bad precision");
  if (!d->h_sigm || !d->h_kv || !d->h_kd || !d->h_ktt || !d->h_kq)
    return fail(LBM_ERR_VALUE, "sigm and damping arrays are required");
  if (d->inistp < 0 || d->inistp > 8) return fail(LBM_ERR_VALUE, "inistp out of range");
  // the Legendre analysis addresses one m of the Fourier columns with 32-bit
  // byte offsets: the largest (9L+2 analysed columns per member) must stay
  // below 4 GiB
  {
    const size_t cx_bytes = d->precision == LBM_FP64 ? 16 : 8;
    const size_t slab = (size_t)d->nlat * (((size_t)(9 * d->nlev + 2) * d->nmem + 7) & ~(size_t)7) * cx_bytes;
    if (slab >= ((size_t)1 << 32))
      return fail(LBM_ERR_VALUE, "nmem=%d: one wavenumber of the analysed Fourier columns (%zu bytes) "
                  "exceeds 4 GiB; use fewer members per plan", d->nmem, slab);
  }
  lbm_plan *p = new lbm_plan();
  p->J = d->nlat;
  p->I = d->nlon;
  p->N = d->nmax;
  p->M = d->nmax + 1;
  p->L = d->nlev;
  p->E = d->nmem;
  p->nb = d->nbasic;
  p->prec = d->precision;
  p->tune[4] = builtin_grid_choice(p);
  p->split = d->precision == LBM_BF16X3;
  p->ffac = ffac;
  p->nspec = p->M * (p->M + 1) / 2;
  p->Jh = p->J / 2;
  p->inistp = d->inistp;
  // one matrix set per start-up dt, one for delt and one for a free dt
  // (Integrator.step with any dt, lbm_prepare_dt)
  p->nset = p->inistp + 2;
  p->user_dt = d->delt;
  p->strictly_linear = d->strictly_linear;
  p->linear = d->tangent_linear ? 1 : 0;
  p->skip_low_n = d->skip_low_n_vdiff;
  p->a = d->radius;
  p->delt = d->delt;
  p->robert = d->robert;
  const int J = p->J, I = p->I, M = p->M, L = p->L, E = p->E;
  int rc = LBM_OK;
  auto bail = [&](int r) {
    lbm_plan_destroy(p);
    return r;
  };
  gauss_legendre(J, p->mu, p->gw);
  p->coslat.resize(J);
  p->lat.resize(J);
  for (int j = 0; j < J; ++j) {
    p->coslat[j] = std::sqrt(1.0 - p->mu[j] * p->mu[j]);
    p->lat[j] = std::asin(p->mu[j]);
  }
  p->lon.resize(I);
  for (int i = 0; i < I; ++i) p->lon[i] = i * (2.0 * M_PI / I);
  if ((rc = build_vertical(p, d->h_sigm)) != LBM_OK) return bail(rc);
  p->kv.assign(d->h_kv, d->h_kv + (size_t)M * L);
  p->kd.assign(d->h_kd, d->h_kd + (size_t)M * L);
  p->ktt.assign(d->h_ktt, d->h_ktt + (size_t)M * L);
  p->kq.assign(d->h_kq, d->h_kq + (size_t)M * L);
  const double *vd[4] = {d->h_vdif_vor, d->h_vdif_div, d->h_vdif_t, d->h_vdif_q};
  for (int f = 0; f < 4; ++f)
    if (vd[f]) p->vdif[f].assign(vd[f], vd[f] + L);
  if ((rc = build_vdiff(p)) != LBM_OK) return bail(rc);
  p->n_of_s.resize(p->nspec);
  for (int m = 0; m < M; ++m)
    for (int n = m; n < M; ++n) p->n_of_s[spec_off(m, M) + n - m] = (short)n;

  auto dalloc = [&](void **ptr, size_t bytes) -> int {
    HIPCHK(hipMalloc(ptr, std::max<size_t>(bytes, 16)));
    return LBM_OK;
  };
  if (hipStreamCreate(&p->own) != hipSuccess) return bail(fail(LBM_ERR_RUNTIME, "stream create failed"));
  // default priority, like every stream of the library
  if (hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(LBM_ERR_RUNTIME, "stream create failed"));
  for (auto &e : p->fj)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
      return bail(fail(LBM_ERR_RUNTIME, "event create failed"));
  if ((rc = build_tabs<double>(p)) != LBM_OK) return bail(rc);
  if (p->prec != LBM_FP64 && (rc = build_tabs<float>(p)) != LBM_OK) return bail(rc);
  if ((rc = set_twiddles<double>()) != LBM_OK) return bail(rc);
  if ((rc = set_twiddles<float>()) != LBM_OK) return bail(rc);
  // identity-free defaults for the reference temperature until a basic state
  p->tbar.assign((size_t)p->nb * L, 0.0);
  const size_t rs = p->rsize(), cs = 2 * rs;
  const size_t EL = (size_t)E * L;
  const ColMap cin{E, L, 6, 6}, cout{E, L, 9, 2};
  struct A {
    void **ptr;
    size_t bytes;
  } allocs[] = {
      {(void **)&p->d_n_of_s, sizeof(short) * p->nspec},
      {(void **)&p->d_mu, sizeof(double) * J},
      {(void **)&p->d_psbar, sizeof(double) * p->nb * J * I},
      {&p->sa3, cs * p->nspec * 4 * EL},
      {&p->sb3, cs * p->nspec * 4 * EL},
      {&p->sa2, cs * p->nspec * 3 * E},
      {&p->tend3, cs * p->nspec * 4 * EL},
      {&p->tend2, cs * p->nspec * 2 * E},
      {&p->fin, cs * M * J * fpad(cin.ncols())},
      {&p->gin, rs * (size_t)E * J * cin.rows() * I},
      {&p->gout, rs * (size_t)E * J * cout.rows() * I},
      {&p->fout, cs * M * J * fpad(cout.ncols())},
      {&p->raw, rs * (size_t)E * J * 4 * L * I},
      {&p->psa, rs * (size_t)E * J * I},
      {&p->psb, rs * (size_t)E * J * I},
      {&p->bsg, rs * (size_t)p->nb * J * BRow{L}.rows() * I},
      {(void **)&p->ctl, sizeof(Ctl)},
      {(void **)&p->pend, sizeof(int) * E},
      {(void **)&p->d_nsteps, sizeof(int) * E},
  };
  for (auto &al : allocs)
    if ((rc = dalloc(al.ptr, al.bytes)) != LBM_OK) return bail(rc);
  if (p->linear) {
    if ((rc = dalloc(&p->sb2, cs * p->nspec * 3 * E)) != LBM_OK) return bail(rc);
    if (hipMemset(p->sb2, 0, cs * p->nspec * 3 * E) != hipSuccess)
      return bail(fail(LBM_ERR_RUNTIME, "initialisation failed"));
  }
  if (hipMemcpy(p->d_n_of_s, p->n_of_s.data(), sizeof(short) * p->nspec, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(p->d_mu, p->mu.data(), sizeof(double) * J, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(LBM_ERR_RUNTIME, "table upload failed"));
  Ctl c0{0, 0, INT_MAX, 0};
  if (hipMemcpy(p->ctl, &c0, sizeof(Ctl), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(p->pend, 0, sizeof(int) * E) != hipSuccess ||
      hipMemset(p->sa3, 0, cs * p->nspec * 4 * EL) != hipSuccess ||
      hipMemset(p->sb3, 0, cs * p->nspec * 4 * EL) != hipSuccess ||
      hipMemset(p->sa2, 0, cs * p->nspec * 3 * E) != hipSuccess ||
      hipMemset(p->psa, 0, rs * (size_t)E * J * I) != hipSuccess ||
      hipMemset(p->psb, 0, rs * (size_t)E * J * I) != hipSuccess)
    return bail(fail(LBM_ERR_RUNTIME, "initialisation failed"));
  p->nsteps_h.assign(E, 0);
  *out = p;
  return LBM_OK;
}

int lbm_plan_get(const lbm_plan *p, const char *name, double *h_out, int64_t n) {
  if (!p || !name || !h_out) return fail(LBM_ERR_VALUE, "null argument");
  const std::vector<double> *v = nullptr;
  std::string s(name);
  if (s == "mu") v = &p->mu;
  else if (s == "gw") v = &p->gw;
  else if (s == "coslat") v = &p->coslat;
  else if (s == "lat") v = &p->lat;
  else if (s == "lon") v = &p->lon;
  else if (s == "sig") v = &p->sig;
  else if (s == "sigm") v = &p->sigm;
  else if (s == "dsig") v = &p->dsig;
  else if (s == "alpha") v = &p->alpha;
  else if (s == "cpg") v = &p->cpg;
  else if (s == "w3") v = &p->w3;
  else if (s == "wP") v = &p->wP;
  else if (s == "dsigm") v = &p->dsigm;
  else if (s == "dlog3") v = &p->dlog3;
  else if (s == "ww") v = &p->ww;
  else if (s == "lop") v = &p->lop;
  else if (s == "tbar") v = &p->tbar;
  else if (s == "minv") v = &p->minv;
  else if (s == "vinv") v = &p->vinv;
  else if (s == "mmat") v = &p->mmat;
  if (!v) return fail(LBM_ERR_VALUE, "unknown plan array '%s'", name);
  const int64_t cnt = (int64_t)v->size();
  if (n < cnt) return fail(LBM_ERR_VALUE, "'%s' needs %lld doubles", name, (long long)cnt);
  std::copy(v->begin(), v->end(), h_out);
  return (int)cnt;
}

int lbm_grid_to_wave(lbm_plan *p, const double *g, double *w, int B, int nk, void *s) {
  return spectral_dispatch(p, OP_G2W, g, nullptr, w, nullptr, B, nk, s);
}
int lbm_wave_to_grid(lbm_plan *p, const double *w, double *g, int B, int nk, void *s) {
  return spectral_dispatch(p, OP_W2G, w, nullptr, g, nullptr, B, nk, s);
}
int lbm_uv_to_vordiv(lbm_plan *p, const double *u, const double *v, double *vor, double *div, int B,
                     int nk, void *s) {
  return spectral_dispatch(p, OP_UV2VD, u, v, vor, div, B, nk, s);
}
int lbm_psichi_to_uv(lbm_plan *p, const double *psi, const double *chi, double *u, double *v, int B,
                     int nk, void *s) {
  return spectral_dispatch(p, OP_PC2UV, psi, chi, u, v, B, nk, s);
}
int lbm_hgrad(lbm_plan *p, const double *w, double *gx, double *gy, int B, int nk, void *s) {
  return spectral_dispatch(p, OP_HGRAD, w, nullptr, gx, gy, B, nk, s);
}
int lbm_grid_to_wave_adj(lbm_plan *p, const double *w, double *g, int B, int nk, void *s) {
  return spectral_dispatch(p, OP_G2W_ADJ, w, nullptr, g, nullptr, B, nk, s);
}
int lbm_wave_to_grid_adj(lbm_plan *p, const double *g, double *w, int B, int nk, void *s) {
  return spectral_dispatch(p, OP_W2G_ADJ, g, nullptr, w, nullptr, B, nk, s);
}
int lbm_uv_to_vordiv_adj(lbm_plan *p, const double *vor, const double *div, double *u, double *v,
                         int B, int nk, void *s) {
  return spectral_dispatch(p, OP_UV2VD_ADJ, vor, div, u, v, B, nk, s);
}
int lbm_psichi_to_uv_adj(lbm_plan *p, const double *u, const double *v, double *psi, double *chi,
                         int B, int nk, void *s) {
  return spectral_dispatch(p, OP_PC2UV_ADJ, u, v, psi, chi, B, nk, s);
}
int lbm_hgrad_adj(lbm_plan *p, const double *gx, const double *gy, double *w, int B, int nk,
                  void *s) {
  return spectral_dispatch(p, OP_HGRAD_ADJ, gx, gy, w, nullptr, B, nk, s);
}

int lbm_set_basic_state(lbm_plan *p, const double *u, const double *v, const double *t,
                        const double *q, const double *ps, void *stream) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  if (!u || !v || !t || !q || !ps) return fail(LBM_ERR_VALUE, "basic state needs u, v, t, q, ps");
  destroy_graphs(p);
  return set_basic(p, u, v, t, q, ps, p->strm(stream));
}

int lbm_set_forcing(lbm_plan *p, const double *frc, int nfrc, void *stream) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  hipStream_t st = p->strm(stream);
  if (p->in_run) p->frc_in_run = true;
  // the captured graphs stay valid while the forcing buffer keeps its
  // address (a new forcing of the same member count between steps)
  if (p->mjo) {
    // a constant forcing replaces the MJO forcing
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipFree(p->mjo));
    p->mjo = nullptr;
    destroy_graphs(p);
  }
  if (!frc) {
    if (p->frc) {
      HIPCHK(hipStreamSynchronize(st));
      HIPCHK(hipFree(p->frc));
      destroy_graphs(p);
    }
    p->frc = nullptr;
    p->nfrc = 0;
    return LBM_OK;
  }
  if (nfrc != 1 && nfrc != p->E) return fail(LBM_ERR_VALUE, "forcing must have 1 or nmem members");
  const int J = p->J, I = p->I, L = p->L;
  const size_t n = (size_t)nfrc * J * I * L;
  if (p->frc && p->nfrc != nfrc) {
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipFree(p->frc));
    p->frc = nullptr;
  }
  if (!p->frc) {
    HIPCHK(hipMalloc(&p->frc, n * p->rsize()));
    destroy_graphs(p);
  }
  // a graph of the previous run may still read the buffer
  if (p->own != st) HIPCHK(hipStreamSynchronize(p->own));
  p->nfrc = nfrc;
  if (p->prec == LBM_FP64)
    hipLaunchKernelGGL((k_pack_grid<double>), dim3(nblk(n)), dim3(256), 0, st, frc, (double *)p->frc,
                       nfrc, J, I, L, L, 0, (const double *)nullptr);
  else
    hipLaunchKernelGGL((k_pack_grid<float>), dim3(nblk(n)), dim3(256), 0, st, frc, (float *)p->frc,
                       nfrc, J, I, L, L, 0, (const double *)nullptr);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

int lbm_set_state(lbm_plan *p, const double *u, const double *v, const double *t, const double *q,
                  const double *ps, void *stream) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  if (!p->basic_set) return fail(LBM_ERR_STATE, "set the basic state before the perturbation");
  hipStream_t st = p->strm(stream);
  p->steps_total = 0;
  p->in_run = false;
  p->levels_equal = true;
  Ctl c0{0, 0, INT_MAX, 0};
  int rc = p->prec == LBM_FP64 ? set_state_t<double>(p, u, v, t, q, ps, st)
                               : set_state_t<float>(p, u, v, t, q, ps, st);
  if (rc != LBM_OK) return rc;
  c0.raw = p->have_raw ? 1 : 0;
  HIPCHK(hipMemcpyAsync(p->ctl, &c0, sizeof(Ctl), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  return LBM_OK;
}

int lbm_set_state_keep(lbm_plan *p, const double *u, const double *v, const double *t,
                       const double *q, const double *ps, void *stream) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  if (!p->basic_set) return fail(LBM_ERR_STATE, "set the basic state before the perturbation");
  if (p->in_run) return fail(LBM_ERR_STATE, "a run is in progress (finish it first)");
  if (!(u || v || t || q || ps)) return LBM_OK;
  hipStream_t st = p->strm(stream);
  const double *in[5] = {u, v, t, q, ps};
  const bool equal = p->steps_total == 0 && p->levels_equal;
  int rc = p->prec == LBM_FP64 ? set_state_keep_t<double>(p, in, st) : set_state_keep_t<float>(p, in, st);
  if (rc != LBM_OK) return rc;
  p->levels_equal = equal;
  p->steps_total = 0;
  Ctl c0{0, p->have_raw ? 1 : 0, INT_MAX, 0};
  HIPCHK(hipMemcpyAsync(p->ctl, &c0, sizeof(Ctl), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  return LBM_OK;
}

int lbm_get_prev_state(lbm_plan *p, double *u, double *v, double *t, double *q, double *ps,
                       void *stream) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  if (!p->basic_set) return fail(LBM_ERR_STATE, "no basic state");
  if (p->in_run) return fail(LBM_ERR_STATE, "a run is in progress (finish it first)");
  hipStream_t st = p->strm(stream);
  return p->prec == LBM_FP64 ? get_level_t<double>(p, 1, u, v, t, q, ps, st)
                             : get_level_t<float>(p, 1, u, v, t, q, ps, st);
}

int lbm_tendencies(lbm_plan *p, const double *u, const double *v, const double *t, const double *q,
                   const double *ps, double *tvor, double *tdiv, double *tt, double *tq,
                   double *tpi, void *stream) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  if (!p->basic_set) return fail(LBM_ERR_STATE, "lbm_tendencies before lbm_set_basic_state");
  if (p->in_run) return fail(LBM_ERR_STATE, "a run is in progress (finish it first)");
  if (p->linear) return fail(LBM_ERR_STATE, "the tangent-linear model has no grid-state tendencies");
  hipStream_t st = p->strm(stream);
  double *out[5] = {tvor, tdiv, tt, tq, tpi};
  return p->prec == LBM_FP64 ? tendencies_t<double>(p, u, v, t, q, ps, out, st)
                             : tendencies_t<float>(p, u, v, t, q, ps, out, st);
}

int lbm_prepare_dt(lbm_plan *p, double dt) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  return prepare_dt(p, dt);
}

int lbm_step(lbm_plan *p, double dt, const double *pvor, const double *pdiv, const double *pt,
             const double *pq, const double *ppi, const double *tvor, const double *tdiv,
             const double *tt, const double *tq, const double *tpi, double *nvor, double *ndiv,
             double *nt, double *nq, double *npi, void *stream) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  if (!p->basic_set) return fail(LBM_ERR_STATE, "lbm_step before lbm_set_basic_state");
  if (p->in_run) return fail(LBM_ERR_STATE, "a run is in progress (finish it first)");
  const double *prev[5] = {pvor, pdiv, pt, pq, ppi}, *tend[5] = {tvor, tdiv, tt, tq, tpi};
  for (int f = 0; f < 5; ++f)
    if (!prev[f] || !tend[f]) return fail(LBM_ERR_VALUE, "lbm_step needs every previous field and tendency");
  hipStream_t st = p->strm(stream);
  double *out[5] = {nvor, ndiv, nt, nq, npi};
  return p->prec == LBM_FP64 ? step_t<double>(p, dt, prev, tend, out, st)
                             : step_t<float>(p, dt, prev, tend, out, st);
}

int lbm_robert_asselin(int64_t n, double *prev, const double *cur, const double *nxt, double gamma,
                       void *stream) {
  if (n < 0 || (n > 0 && (!prev || !cur || !nxt))) return fail(LBM_ERR_VALUE, "bad arguments");
  if (n == 0) return LBM_OK;
  hipLaunchKernelGGL(k_robert_asselin, dim3(nblk((size_t)n)), dim3(256), 0, (hipStream_t)stream, prev,
                     cur, nxt, gamma, (size_t)n);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

int lbm_lnps(int64_t n, int64_t nper, const double *in, const double *psbar, double *out,
             int inverse, void *stream) {
  if (n < 0 || nper <= 0 || n % nper || (n > 0 && (!in || !psbar || !out)))
    return fail(LBM_ERR_VALUE, "bad arguments (n must be a multiple of nper)");
  if (n == 0) return LBM_OK;
  hipLaunchKernelGGL(k_lnps, dim3(nblk((size_t)n)), dim3(256), 0, (hipStream_t)stream, in, psbar, out,
                     (size_t)n, (size_t)nper, inverse);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

int lbm_grid_stats(int64_t n, const double *x, double *h_out, void *stream) {
  if (!h_out || n < 0 || (n > 0 && !x)) return fail(LBM_ERR_VALUE, "bad arguments");
  hipStream_t st = (hipStream_t)stream;
  unsigned long long *d = nullptr, h[2] = {0, 0};
  HIPCHK(hipMalloc((void **)&d, sizeof(h)));
  int rc = LBM_OK;
  if (hipMemsetAsync(d, 0, sizeof(h), st) != hipSuccess) rc = fail(LBM_ERR_RUNTIME, "memset failed");
  if (rc == LBM_OK && n > 0) {
    const unsigned nb = (unsigned)std::min<size_t>(((size_t)n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_grid_stats, dim3(nb), dim3(256), 0, st, x, (size_t)n, d);
    if (hipGetLastError() != hipSuccess) rc = fail(LBM_ERR_RUNTIME, "launch failed");
  }
  if (rc == LBM_OK && (hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, st) != hipSuccess ||
                       hipStreamSynchronize(st) != hipSuccess))
    rc = fail(LBM_ERR_RUNTIME, "statistics copy failed");
  (void)hipFree(d);
  if (rc != LBM_OK) return rc;
  double mx;
  std::memcpy(&mx, &h[0], sizeof(double));
  h_out[0] = mx;
  h_out[1] = (double)h[1];
  return LBM_OK;
}

int lbm_grid_stats_noalloc(int64_t n, const double *x, double *h_out, void *stream) {
  if (!h_out || n < 0 || (n > 0 && !x)) return fail(LBM_ERR_VALUE, "bad arguments");
  unsigned long long h[2] = {0, 0};
  if (n > 0) {
    unsigned long long *s = nullptr;
    CHK(device_scratch(&s));
    // the grid-stride launch never has more blocks than scratch has pairs,
    // so every n fits; each block writes its own pair, nothing is cleared
    const size_t nb = std::min<size_t>(((size_t)n + 255) / 256, kStatsSlots);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_grid_stats_part, dim3((unsigned)nb), dim3(256), 0, st, x, (size_t)n,
                       s + kStatsOff);
    hipLaunchKernelGGL(k_grid_stats_final, dim3(1), dim3(256), 0, st,
                       (const unsigned long long *)(s + kStatsOff), (int)nb, s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h, s, sizeof(h), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  double mx;
  std::memcpy(&mx, &h[0], sizeof(double));
  h_out[0] = mx;
  h_out[1] = (double)h[1];
  return LBM_OK;
}

int lbm_spec_scale(lbm_plan *p, const double *in, double *out, int B, int nk, int kind,
                   void *stream) {
  if (!p || !in || !out || B < 0 || nk < 1) return fail(LBM_ERR_VALUE, "bad arguments");
  if (kind != 0 && kind != 1) return fail(LBM_ERR_VALUE, "kind must be 0 (laplacian) or 1 (inverse)");
  const size_t n = (size_t)B * p->M * p->M * nk * 2;
  if (n == 0) return LBM_OK;
  hipLaunchKernelGGL(k_spec_scale, dim3(nblk(n)), dim3(256), 0, p->strm(stream), in, out,
                     kind == 0 ? p->d64.lap : p->d64.inv_lap, n, p->M, nk);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

int lbm_hydrostatic(const double *in, double *out, int64_t npts, int nlev, int inner,
                    const double *alpha, const double *dlog3, void *stream) {
  if (!in || !out || !alpha || !dlog3 || npts < 0 || nlev < 1 || (inner != 1 && inner != 2))
    return fail(LBM_ERR_VALUE, "bad arguments (inner is 1 for real, 2 for complex)");
  if (npts == 0) return LBM_OK;
  hipLaunchKernelGGL(k_hydrostatic, dim3(nblk((size_t)npts * inner)), dim3(256), 0, (hipStream_t)stream,
                     in, out, alpha, dlog3, (size_t)npts, nlev, inner);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

int lbm_autotune(lbm_plan *p, int reps, char *report, int64_t len) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  if (reps < 1) return fail(LBM_ERR_VALUE, "reps must be >= 1");
  if (!p->basic_set) return fail(LBM_ERR_STATE, "autotune after lbm_set_basic_state");
  if (p->in_run) return fail(LBM_ERR_STATE, "a run is in progress (finish it first)");
  if (p->prof) return fail(LBM_ERR_STATE, "stop profiling before autotuning");
  destroy_graphs(p);
  std::string rep;
  const int rc = p->prec == LBM_FP64 ? autotune_t<double>(p, reps, rep) : autotune_t<float>(p, reps, rep);
  if (rc != LBM_OK) return rc;
  if (report && len > 0) std::snprintf(report, (size_t)len, "%s", rep.c_str());
  return LBM_OK;
}

int lbm_set_tuning(lbm_plan *p, const int32_t *h_choice, int n) {
  if (!p || (n > 0 && !h_choice)) return fail(LBM_ERR_VALUE, "null argument");
  if (n < 0 || n > lbm_plan::NTUNE) return fail(LBM_ERR_VALUE, "at most %d choices", lbm_plan::NTUNE);
  for (int k = 0; k < n; ++k)
    if (h_choice[k] < 0 || h_choice[k] > (k == 4 ? 4 : 2))
      return fail(LBM_ERR_VALUE, "choice %d out of range", k);
  if (p->in_run) return fail(LBM_ERR_STATE, "a run is in progress (finish it first)");
  destroy_graphs(p);
  for (int k = 0; k < lbm_plan::NTUNE; ++k) p->tune[k] = k < n ? h_choice[k] : 0;
  return LBM_OK;
}

int lbm_get_tuning(const lbm_plan *p, int32_t *h_choice, int n) {
  if (!p || !h_choice || n < lbm_plan::NTUNE) return fail(LBM_ERR_VALUE, "need %d entries", lbm_plan::NTUNE);
  for (int k = 0; k < lbm_plan::NTUNE; ++k) h_choice[k] = p->tune[k];
  return lbm_plan::NTUNE;
}

int lbm_set_checkpoints(lbm_plan *p, int every) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  if (every < 0) return fail(LBM_ERR_VALUE, "every must be >= 0");
  if (p->in_run) return fail(LBM_ERR_STATE, "a run is in progress (finish it first)");
  p->ckpt_every = every;
  p->ckpt_ok = false;
  if (every == 0) {
    void **bufs[] = {&p->ckpt, &p->traj, &p->keep, &p->psa_a, &p->psb_a};
    HIPCHK(hipDeviceSynchronize());
    for (void **b : bufs) {
      if (*b) HIPCHK(hipFree(*b));
      *b = nullptr;
    }
    p->ckpt_cap = p->traj_cap = 0;
  }
  return LBM_OK;
}

int lbm_run_abort(lbm_plan *p) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  p->in_run = false;
  return LBM_OK;
}

int lbm_set_lengths(lbm_plan *p, const int32_t *h_nsteps, void *stream) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  (void)stream;
  if (!h_nsteps) {
    std::fill(p->nsteps_h.begin(), p->nsteps_h.end(), 0);
    return LBM_OK;
  }
  for (int e = 0; e < p->E; ++e) {
    if (h_nsteps[e] < 0) return fail(LBM_ERR_VALUE, "negative length for member %d", e);
    p->nsteps_h[e] = h_nsteps[e];
  }
  return LBM_OK;
}

// The run loop behind lbm_run, lbm_run_to and lbm_run_steps: runs (or
// continues) the run of nsteps up to full step stop (0: its end), or, with
// count > 0, by count more executed steps (start-up steps included).
int run_core(lbm_plan *p, int nsteps, int stop, int count, int use_graph, int check_every,
             void *stream) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  if (!p->basic_set) return fail(LBM_ERR_STATE, "lbm_run before lbm_set_basic_state");
  hipStream_t st = use_graph ? p->own : p->strm(stream);
  if (use_graph && stream) HIPCHK(hipStreamSynchronize((hipStream_t)stream));
  std::vector<int> len(p->E);
  for (int e = 0; e < p->E; ++e) len[e] = nsteps > 0 ? nsteps : p->nsteps_h[e];
  const int maxlen = *std::max_element(len.begin(), len.end());
  const int total = p->inistp + maxlen;
  if (maxlen <= 0) {
    p->in_run = false;
    return LBM_OK;
  }
  if (stop < 0 || stop > maxlen) return fail(LBM_ERR_VALUE, "stop=%d outside 0..%d", stop, maxlen);
  const int target_stop = stop == 0 ? total : p->inistp + stop;
  if (p->in_run && len != p->run_len)
    return fail(LBM_ERR_STATE, "a run of other lengths is in progress (finish it or set a state)");
  if (p->in_run && target_stop < p->run_done)
    return fail(LBM_ERR_VALUE, "stop=%d is behind the %d full steps already run", stop,
                p->run_done - p->inistp);
  if (!p->in_run) {
    HIPCHK(hipMemcpyAsync(p->d_nsteps, len.data(), sizeof(int) * p->E, hipMemcpyHostToDevice, st));
    // every run restarts the start-up half steps (model.py:211-219)
    Ctl h;
    HIPCHK(hipMemcpyAsync(&h, p->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    h.step = 0;
    h.bad_step = INT_MAX;
    HIPCHK(hipMemcpyAsync(p->ctl, &h, sizeof(Ctl), hipMemcpyHostToDevice, st));
    p->in_run = true;
    p->frc_in_run = false;
    p->run_len = len;
    p->run_done = 0;
    // checkpoints of this run for the adjoint of the default model
    p->ckpt_ok = false;
    if (p->ckpt_every > 0 && !p->linear) {
      p->ckpt_count = (total + p->ckpt_every - 1) / p->ckpt_every;
      CHK(grow(&p->ckpt, &p->ckpt_cap, (size_t)p->ckpt_count * state_parts(p).total, st));
      p->ckpt_len = len;
      p->ckpt_total = total;
      p->ckpt_raw = p->raw_pending;
    }
  }
  const bool ckpt_on = p->ckpt_every > 0 && !p->linear;
  auto ckpt_at = [&](int step) -> int {
    if (ckpt_on && step % p->ckpt_every == 0)
      CHK(copy_state(p, (char *)p->ckpt + (size_t)(step / p->ckpt_every) * state_parts(p).total, true, st));
    return LBM_OK;
  };
  const int target = count > 0 ? std::min(total, p->run_done + count) : target_stop;
  const int chunk = 8;
  int done = p->run_done;
  int rc = LBM_OK;
  if (done == 0 && p->raw_pending && done < target) {
    // the first step after a raw initial state reads the raw grid (unfused)
    CHK(ckpt_at(0));
    CHK(step_dispatch(p, st, all_live(p, len, 0) ? 0 : live_members(p, len, 0, true), true));
    p->raw_pending = false;
    done = 1;
  }
  while (done < target) {
    int todo = target - done;
    if (check_every > 0) todo = std::min(todo, check_every - (done % check_every));
    if (ckpt_on) {
      CHK(ckpt_at(done));
      todo = std::min(todo, p->ckpt_every - done % p->ckpt_every);
    }
    const int n = (use_graph && todo < chunk) ? 1 : std::min(todo, chunk);
    // the members still live at the first step of this stretch cover it;
    // no mask while every member advances through it
    const int ecnt = all_live(p, len, done + n - 1) ? 0 : live_members(p, len, done, true);
    if (use_graph) {
      hipGraphExec_t ge = nullptr;
      CHK(capture_graph(p, n == chunk ? 0 : 1, n, ecnt, st, &ge));
      HIPCHK(hipGraphLaunch(ge, st));
    } else {
      for (int s = 0; s < n; ++s) CHK(step_dispatch(p, st, ecnt));
    }
    done += n;
    if (check_every > 0 && done % check_every == 0 && (rc = check_bad(p, st, done)) != LBM_OK) break;
  }
  p->steps_total += done - p->run_done;
  p->run_done = done;
  if (ckpt_on && rc == LBM_OK && done == total) p->ckpt_ok = true;
  if (rc != LBM_OK || done == total) p->in_run = false;
  if (rc != LBM_OK) return rc;
  if (done == total) CHK(p->prec == LBM_FP64 ? final_check<double>(p, st) : final_check<float>(p, st));
  CHK(check_bad(p, st, done));
  if (use_graph && stream) HIPCHK(hipStreamSynchronize(st));
  return LBM_OK;
}

int lbm_run_to(lbm_plan *p, int nsteps, int stop, int use_graph, int check_every, void *stream) {
  return run_core(p, nsteps, stop, 0, use_graph, check_every, stream);
}

int lbm_run_steps(lbm_plan *p, int nsteps, int count, int use_graph, int check_every,
                  void *stream) {
  if (count < 1) return fail(LBM_ERR_VALUE, "count must be >= 1");
  return run_core(p, nsteps, 0, count, use_graph, check_every, stream);
}

int lbm_run(lbm_plan *p, int nsteps, int use_graph, int check_every, void *stream) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  p->in_run = false;
  return lbm_run_to(p, nsteps, 0, use_graph, check_every, stream);
}

int lbm_get_state(lbm_plan *p, double *u, double *v, double *t, double *q, double *ps,
                  void *stream) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  if (!p->basic_set) return fail(LBM_ERR_STATE, "no basic state");
  hipStream_t st = p->strm(stream);
  return p->prec == LBM_FP64 ? get_state_t<double>(p, u, v, t, q, ps, st)
                             : get_state_t<float>(p, u, v, t, q, ps, st);
}

int lbm_get_spectral(lbm_plan *p, int level, double *vor, double *div, double *t, double *q,
                     double *pi, void *stream) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  if (level != 0 && level != 1) return fail(LBM_ERR_VALUE, "level must be 0 or 1");
  hipStream_t st = p->strm(stream);
  const int E = p->E, M = p->M, L = p->L;
  double *dst[4] = {vor, div, t, q};
  const void *src3 = level == 0 ? p->sa3 : p->sb3;
  const size_t tot = (size_t)E * M * M * L;
  for (int f = 0; f < 4; ++f) {
    if (!dst[f]) continue;
    if (p->prec == LBM_FP64)
      hipLaunchKernelGGL((k_unpack_spec<double>), dim3(nblk(tot)), dim3(256), 0, st,
                         (const cx<double> *)src3, dst[f], E, M, L, 4, f, E * L);
    else
      hipLaunchKernelGGL((k_unpack_spec<float>), dim3(nblk(tot)), dim3(256), 0, st,
                         (const cx<float> *)src3, dst[f], E, M, L, 4, f, E * L);
  }
  if (pi && level == 0) {
    if (p->prec == LBM_FP64)
      hipLaunchKernelGGL((k_unpack_spec<double>), dim3(nblk((size_t)E * M * M)), dim3(256), 0, st,
                         (const cx<double> *)p->sa2, pi, E, M, 1, 3, 0, E);
    else
      hipLaunchKernelGGL((k_unpack_spec<float>), dim3(nblk((size_t)E * M * M)), dim3(256), 0, st,
                         (const cx<float> *)p->sa2, pi, E, M, 1, 3, 0, E);
  }
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

int lbm_profile(lbm_plan *p, int enable) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  if (enable && !p->ev[0])
    for (auto &e : p->ev) HIPCHK(hipEventCreate(&e));
  p->prof = enable != 0;
  return LBM_OK;
}

int lbm_profile_read(lbm_plan *p, double *h_ms, int64_t *h_count, int n) {
  if (!p || !h_ms || !h_count) return fail(LBM_ERR_VALUE, "null argument");
  const int k = std::min(n, lbm_plan::NPROF);
  for (int i = 0; i < k; ++i) {
    h_ms[i] = p->prof_ms[i];
    h_count[i] = p->prof_n[i];
    p->prof_ms[i] = 0.0;
    p->prof_n[i] = 0;
  }
  return k;
}

int lbm_steps_taken(const lbm_plan *p, int64_t *out) {
  if (!p || !out) return fail(LBM_ERR_VALUE, "null argument");
  *out = p->steps_total;
  return LBM_OK;
}

int lbm_set_spectral_state(lbm_plan *p, const double *vor, const double *div, const double *t,
                           const double *q, const double *pi, void *stream) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  if (!p->basic_set) return fail(LBM_ERR_STATE, "set the basic state before the perturbation");
  hipStream_t st = p->strm(stream);
  p->steps_total = 0;
  p->in_run = false;
  p->levels_equal = true;
  int rc = p->prec == LBM_FP64 ? set_spectral_t<double>(p, vor, div, t, q, pi, st)
                               : set_spectral_t<float>(p, vor, div, t, q, pi, st);
  if (rc != LBM_OK) return rc;
  if (p->jacc) HIPCHK(hipMemsetAsync(p->jacc, 0, sizeof(double) * p->E, st));
  Ctl c0{0, 0, INT_MAX, 0};
  HIPCHK(hipMemcpyAsync(p->ctl, &c0, sizeof(Ctl), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  return LBM_OK;
}

int lbm_set_response(lbm_plan *p, const double *wgrid, const double *wlev, const double *wstep,
                     int nstep) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  destroy_graphs(p);
  if (!wgrid) {
    p->resp_set = false;
    return LBM_OK;
  }
  if (!wlev || (nstep > 0 && !wstep) || nstep < 0)
    return fail(LBM_ERR_VALUE, "response needs level weights and nstep >= 0 step weights");
  hipStream_t st = p->own;
  return p->prec == LBM_FP64 ? set_response_t<double>(p, wgrid, wlev, wstep, nstep, st)
                             : set_response_t<float>(p, wgrid, wlev, wstep, nstep, st);
}

int lbm_get_response(lbm_plan *p, double *h_j) {
  if (!p || !h_j) return fail(LBM_ERR_VALUE, "null argument");
  if (!p->resp_set) return fail(LBM_ERR_STATE, "no response set");
  HIPCHK(hipStreamSynchronize(p->own));
  HIPCHK(hipMemcpy(h_j, p->jacc, sizeof(double) * p->E, hipMemcpyDeviceToHost));
  return LBM_OK;
}

int lbm_adjoint_run(lbm_plan *p, int nsteps, double *gvor, double *gdiv, double *gt, double *gq,
                    double *gpi, double *gfrc, int use_graph, void *stream) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  if (!p->basic_set) return fail(LBM_ERR_STATE, "lbm_adjoint_run before lbm_set_basic_state");
  if (!p->resp_set) return fail(LBM_ERR_STATE, "the adjoint needs a response (lbm_set_response)");
  if (p->mjo)
    return fail(LBM_ERR_STATE, "the adjoint is built for constant forcings (an MJO forcing is set)");
  if (p->frc_in_run)
    return fail(LBM_ERR_STATE, "the adjoint is built for constant forcings (the forcing changed "
                "between the steps of the last run)");
  if (p->linear && (p->sfv || p->sfp))
    return fail(LBM_ERR_STATE,
                "the adjoint is built for the linear model (steady vorticity, divergence or ln ps "
                "forcing is set)");
  if (!p->linear) {
    // the default model: about the checkpointed trajectory of the last run
    if (p->ckpt_every <= 0 || !p->ckpt_ok)
      return fail(LBM_ERR_STATE, "the adjoint of the default model needs the checkpoints of a "
                  "finished run (lbm_set_checkpoints before lbm_run)");
    if (p->ckpt_raw)
      return fail(LBM_ERR_STATE, "the checkpointed run must start from lbm_set_spectral_state");
    if (p->in_run) return fail(LBM_ERR_STATE, "a run is in progress (finish it first)");
    std::vector<int> len(p->E);
    for (int e = 0; e < p->E; ++e) len[e] = nsteps > 0 ? nsteps : p->nsteps_h[e];
    if (len != p->ckpt_len)
      return fail(LBM_ERR_VALUE, "the lengths differ from the checkpointed run's");
    hipStream_t st = p->strm(stream);
    const size_t EL = (size_t)p->E * p->L, cs = 2 * p->rsize();
    struct A {
      void **ptr;
      size_t bytes;
    } bufs[] = {{&p->aa3, cs * p->nspec * 4 * EL}, {&p->ab3, cs * p->nspec * 4 * EL},
                {&p->aa2, cs * p->nspec * p->E},   {&p->ab2, cs * p->nspec * p->E},
                {&p->s3a, cs * p->nspec * 3 * p->E},
                {&p->frc_a, p->rsize() * (size_t)p->E * p->J * p->I * p->L}};
    for (auto &b : bufs) {
      if (!*b.ptr) HIPCHK(hipMalloc(b.ptr, b.bytes));
      HIPCHK(hipMemsetAsync(*b.ptr, 0, b.bytes, st));
    }
    HIPCHK(hipMemcpyAsync(p->d_nsteps, len.data(), sizeof(int) * p->E, hipMemcpyHostToDevice, st));
    double *out[6] = {gvor, gdiv, gt, gq, gpi, gfrc};
    (void)use_graph;  // the sweep follows each recomputed segment eagerly
    CHK(p->prec == LBM_FP64 ? adjoint_traj_t<double>(p, len, out, st)
                            : adjoint_traj_t<float>(p, len, out, st));
    HIPCHK(hipStreamSynchronize(st));
    return LBM_OK;
  }
  p->in_run = false;
  hipStream_t st = use_graph ? p->own : p->strm(stream);
  if (use_graph && stream) HIPCHK(hipStreamSynchronize((hipStream_t)stream));
  const int E = p->E, L = p->L, J = p->J, I = p->I;
  const size_t EL = (size_t)E * L, cs = 2 * p->rsize();
  std::vector<int> len(E);
  for (int e = 0; e < E; ++e) len[e] = nsteps > 0 ? nsteps : p->nsteps_h[e];
  const int maxlen = *std::max_element(len.begin(), len.end());
  const int total = p->inistp + maxlen;
  struct A {
    void **ptr;
    size_t bytes;
  } bufs[] = {{&p->aa3, cs * p->nspec * 4 * EL}, {&p->ab3, cs * p->nspec * 4 * EL},
              {&p->aa2, cs * p->nspec * E},      {&p->ab2, cs * p->nspec * E},
              {&p->s3a, cs * p->nspec * 3 * E},  {&p->frc_a, p->rsize() * E * J * I * L}};
  for (auto &b : bufs) {
    if (!*b.ptr) HIPCHK(hipMalloc(b.ptr, b.bytes));
    HIPCHK(hipMemsetAsync(*b.ptr, 0, b.bytes, st));
  }
  HIPCHK(hipMemcpyAsync(p->d_nsteps, len.data(), sizeof(int) * E, hipMemcpyHostToDevice, st));
  Ctl saved;
  HIPCHK(hipMemcpyAsync(&saved, p->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  Ctl c0{total - 1, 0, INT_MAX, 0};
  HIPCHK(hipMemcpyAsync(p->ctl, &c0, sizeof(Ctl), hipMemcpyHostToDevice, st));
  const int chunk = 8;
  int done = 0;
  while (done < total) {
    const int todo = total - done;
    const int n = (use_graph && todo < chunk) ? 1 : std::min(todo, chunk);
    // adjoint steps total-1-done down to total-done-n: the earliest has the
    // most live members
    const int ecnt =
        all_live(p, len, total - done - 1) ? 0 : live_members(p, len, total - done - n, false);
    if (use_graph) {
      hipGraphExec_t ge = nullptr;
      CHK(capture_adjoint(p, n == chunk ? 2 : 3, n, ecnt, st, &ge));
      HIPCHK(hipGraphLaunch(ge, st));
    } else {
      for (int s = 0; s < n; ++s) CHK(adjoint_dispatch(p, st, ecnt));
    }
    done += n;
  }
  CHK(p->prec == LBM_FP64 ? adjoint_finish<double>(p, gvor, gdiv, gt, gq, gpi, gfrc, st)
                          : adjoint_finish<float>(p, gvor, gdiv, gt, gq, gpi, gfrc, st));
  // the forward loop control is left as it was
  HIPCHK(hipMemcpyAsync(p->ctl, &saved, sizeof(Ctl), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  return LBM_OK;
}

int lbm_run_segment_done(const lbm_plan *p, int32_t *full_steps) {
  if (!p || !full_steps) return fail(LBM_ERR_VALUE, "null argument");
  *full_steps = p->in_run ? std::max(0, p->run_done - p->inistp) : 0;
  return LBM_OK;
}

int lbm_set_history(lbm_plan *p, int interval_steps, int time_average, int keep_last,
                    void *stream) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  if (interval_steps < 0) return fail(LBM_ERR_VALUE, "interval_steps must be >= 0");
  if (p->in_run) return fail(LBM_ERR_STATE, "history cannot change while a run is in progress");
  hipStream_t st = p->strm(stream);
  destroy_graphs(p);
  const int E = p->E, L = p->L, J = p->J, I = p->I;
  const size_t cs = 2 * p->rsize(), EL = (size_t)E * L;
  const size_t bs = cs * p->nspec * 3 * EL, bg = p->rsize() * (size_t)E * J * (L + 1) * I;
  p->keep_last = keep_last != 0;
  if (p->keep_last && !p->hlast_spec) {
    HIPCHK(hipMalloc(&p->hlast_spec, cs * p->nspec * 2 * EL));
    HIPCHK(hipMalloc(&p->hlast_grid, p->rsize() * (size_t)E * J * L * I));
    HIPCHK(hipMemsetAsync(p->hlast_spec, 0, cs * p->nspec * 2 * EL, st));
    HIPCHK(hipMemsetAsync(p->hlast_grid, 0, p->rsize() * (size_t)E * J * L * I, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  if (interval_steps == 0) {
    HIPCHK(hipStreamSynchronize(st));
    for (int k = 0; k < 2; ++k) {
      if (p->hspec[k]) HIPCHK(hipFree(p->hspec[k]));
      if (p->hgrid[k]) HIPCHK(hipFree(p->hgrid[k]));
      p->hspec[k] = p->hgrid[k] = nullptr;
    }
    p->hist_mode = 0;
    p->hist_K = 0;
    return LBM_OK;
  }
  for (int k = 0; k < 2; ++k) {
    if (!p->hspec[k]) HIPCHK(hipMalloc(&p->hspec[k], bs));
    if (!p->hgrid[k]) HIPCHK(hipMalloc(&p->hgrid[k], bg));
    HIPCHK(hipMemsetAsync(p->hspec[k], 0, bs, st));
    HIPCHK(hipMemsetAsync(p->hgrid[k], 0, bg, st));
  }
  p->hist_mode = time_average ? 1 : 2;
  p->hist_K = interval_steps;
  HIPCHK(hipStreamSynchronize(st));
  return LBM_OK;
}

int lbm_get_history(lbm_plan *p, int interval, double *psi, double *chi, double *u, double *v,
                    double *w, double *t, double *z, double *ps, void *stream) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  if (p->hist_mode == 0) return fail(LBM_ERR_STATE, "no history set (lbm_set_history)");
  if (interval < 1) return fail(LBM_ERR_VALUE, "intervals are numbered from 1");
  hipStream_t st = p->strm(stream);
  if (p->own != st) HIPCHK(hipStreamSynchronize(p->own));
  const int slot = interval & 1;
  return p->prec == LBM_FP64 ? get_history_t<double>(p, slot, psi, chi, u, v, w, t, z, ps, st)
                             : get_history_t<float>(p, slot, psi, chi, u, v, w, t, z, ps, st);
}

int lbm_get_diagnostics(lbm_plan *p, double *psi, double *chi, double *w, double *z,
                        void *stream) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  if (!p->basic_set) return fail(LBM_ERR_STATE, "no basic state");
  if (!p->hlast_spec)
    return fail(LBM_ERR_STATE, "the last top-of-step sample is not kept (lbm_set_history keep_last)");
  hipStream_t st = p->strm(stream);
  if (p->own != st) HIPCHK(hipStreamSynchronize(p->own));
  return p->prec == LBM_FP64 ? get_diagnostics_t<double>(p, psi, chi, w, z, st)
                             : get_diagnostics_t<float>(p, psi, chi, w, z, st);
}

int lbm_set_mjo_forcing(lbm_plan *p, const double *gtf, int nf, int ndays, int nt, int spinup,
                        int call0, int single, void *stream) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  if (!gtf) return fail(LBM_ERR_VALUE, "null forcing");
  if (nf != 1 && nf != p->E) return fail(LBM_ERR_VALUE, "the forcing must have 1 or nmem members");
  if (ndays < 1 || nt < 1 || spinup < 0 || call0 < 0)
    return fail(LBM_ERR_VALUE, "ndays, nt >= 1 and spinup, call0 >= 0 are needed");
  hipStream_t st = p->strm(stream);
  destroy_graphs(p);
  const int J = p->J, I = p->I, L = p->L;
  const size_t n = (size_t)J * L * I, rs = p->rsize();
  HIPCHK(hipStreamSynchronize(st));
  if (p->mjo) HIPCHK(hipFree(p->mjo));
  p->mjo = nullptr;
  if (p->frc && p->nfrc != nf) {
    HIPCHK(hipFree(p->frc));
    p->frc = nullptr;
  }
  if (!p->frc) HIPCHK(hipMalloc(&p->frc, (size_t)nf * n * rs));
  p->nfrc = nf;
  HIPCHK(hipMalloc(&p->mjo, (size_t)nf * ndays * n * rs));
  // (nf * ndays, J, I, L) float64 -> [f][day][j][L][I]
  const int B = nf * ndays;
  if (p->prec == LBM_FP64)
    hipLaunchKernelGGL((k_pack_grid<double>), dim3(nblk((size_t)B * n)), dim3(256), 0, st, gtf,
                       (double *)p->mjo, B, J, I, L, L, 0, (const double *)nullptr);
  else
    hipLaunchKernelGGL((k_pack_grid<float>), dim3(nblk((size_t)B * n)), dim3(256), 0, st, gtf,
                       (float *)p->mjo, B, J, I, L, L, 0, (const double *)nullptr);
  HIPCHK(hipGetLastError());
  p->mjo_nd = ndays;
  p->mjo_nt = nt;
  p->mjo_spin = spinup;
  p->mjo_call0 = call0;
  p->mjo_single = single ? 1 : 0;
  HIPCHK(hipStreamSynchronize(st));
  return LBM_OK;
}

int lbm_spec_resize(const double *src, double *dst, int B, int src_nmax, int dst_nmax, int nk,
                    void *stream) {
  if (!src || !dst) return fail(LBM_ERR_VALUE, "null array");
  if (B < 1 || src_nmax < 0 || dst_nmax < 0 || nk < 1)
    return fail(LBM_ERR_VALUE, "B, nk >= 1 and truncations >= 0 are needed");
  const size_t tot = (size_t)B * (dst_nmax + 1) * (dst_nmax + 1) * nk;
  hipLaunchKernelGGL(k_spec_resize, dim3(nblk(tot)), dim3(256), 0, (hipStream_t)stream,
                     (const double2 *)src, (double2 *)dst, B, src_nmax + 1, dst_nmax + 1, nk);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

int lbm_set_steady_forcing(lbm_plan *p, const double *vor, const double *div, const double *ps,
                           int nf, void *stream) {
  if (!p) return fail(LBM_ERR_VALUE, "null plan");
  hipStream_t st = p->strm(stream);
  destroy_graphs(p);
  HIPCHK(hipStreamSynchronize(st));
  if (p->sfv) HIPCHK(hipFree(p->sfv));
  if (p->sfp) HIPCHK(hipFree(p->sfp));
  p->sfv = p->sfp = nullptr;
  p->nsf = 0;
  if (!vor && !div && !ps) return LBM_OK;
  if (nf != 1 && nf != p->E) return fail(LBM_ERR_VALUE, "the forcing must have 1 or nmem members");
  p->nsf = nf;
  return p->prec == LBM_FP64 ? set_steady_t<double>(p, vor, div, ps, nf, st)
                             : set_steady_t<float>(p, vor, div, ps, nf, st);
}

}  // extern "C"
