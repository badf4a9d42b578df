// Device code of the gfx950 time-stepping core: complex helpers, the
// length-128 real FFT pair, the per-m Legendre synthesis/analysis with
// north/south folding, the fused grid-point tendency kernel and the fused
// spectral semi-implicit leapfrog + vertical diffusion + Robert-Asselin
// kernel.  Everything is templated on the arithmetic type T (double/float).
//
// Internal layouts (E members, L levels, M = nmax+1, J lats, Jh = J/2):
//   spectral  [s][f][ncols]  complex, s = packed (m, n>=m) index,
//             ncols = E*L for level fields, E for surface fields
//   Fourier   [m][j][C]      complex, C = nF3*E*L + nF2*E columns
//   grid      [e][j][r][I]   real, rows r = f*L + l (level fields) then
//             nF3*L + g (surface fields); longitude fastest
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

template <class T>
struct alignas(2 * sizeof(T)) cx {
  T x, y;
};

template <class T>
__host__ __device__ inline cx<T> mk(T a, T b) { cx<T> r; r.x = a; r.y = b; return r; }
template <class T>
__host__ __device__ inline cx<T> operator+(cx<T> a, cx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <class T>
__host__ __device__ inline cx<T> operator-(cx<T> a, cx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <class T>
__host__ __device__ inline cx<T> operator-(cx<T> a) { return mk(-a.x, -a.y); }
template <class T>
__host__ __device__ inline cx<T> operator*(T s, cx<T> a) { return mk(s * a.x, s * a.y); }
template <class T>
__host__ __device__ inline cx<T> cmul(cx<T> a, cx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <class T>
__host__ __device__ inline cx<T> conj(cx<T> a) { return mk(a.x, -a.y); }
// (i*m) * a
template <class T>
__host__ __device__ inline cx<T> im_mul(T m, cx<T> a) { return mk(-m * a.y, m * a.x); }
template <class T>
__host__ __device__ inline void fma_acc(cx<T> &acc, T s, cx<T> a) {
  acc.x = fma(s, a.x, acc.x);
  acc.y = fma(s, a.y, acc.y);
}
template <class T>
__host__ __device__ inline cx<T> zero() { return mk<T>(0, 0); }

// Packed triangular index of (m, n=m).
__host__ __device__ inline int spec_off(int m, int M) { return m * M - (m * (m - 1)) / 2; }

// ---------------------------------------------------------------------------
// Length-8 DFT in place, natural order in and out.  SIGN = -1 forward
// (exp(-2 pi i nk/8)), +1 backward, unnormalised.  Radix-2 split into two
// length-4 DFTs (even/odd samples).
// ---------------------------------------------------------------------------
template <int SIGN, class T>
__host__ __device__ inline void dft4(cx<T> &a0, cx<T> &a1, cx<T> &a2, cx<T> &a3) {
  // j_s = exp(SIGN * i pi/2) = SIGN * i
  cx<T> s02 = a0 + a2, d02 = a0 - a2, s13 = a1 + a3, d13 = a1 - a3;
  cx<T> jd13 = (SIGN > 0) ? mk(-d13.y, d13.x) : mk(d13.y, -d13.x);
  a0 = s02 + s13;
  a2 = s02 - s13;
  a1 = d02 + jd13;
  a3 = d02 - jd13;
}

template <int SIGN, class T>
__host__ __device__ inline void dft8(cx<T> *v) {
  cx<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  cx<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<SIGN>(e0, e1, e2, e3);
  dft4<SIGN>(o0, o1, o2, o3);
  const T r = (T)0.70710678118654752440;
  // w^k = exp(SIGN * 2 pi i k / 8)
  cx<T> w1o1 = cmul(mk(r, (T)SIGN * r), o1);
  cx<T> w2o2 = (SIGN > 0) ? mk(-o2.y, o2.x) : mk(o2.y, -o2.x);
  cx<T> w3o3 = cmul(mk(-r, (T)SIGN * r), o3);
  v[0] = e0 + o0;
  v[4] = e0 - o0;
  v[1] = e1 + w1o1;
  v[5] = e1 - w1o1;
  v[2] = e2 + w2o2;
  v[6] = e2 - w2o2;
  v[3] = e3 + w3o3;
  v[7] = e3 - w3o3;
}

// Column-space map: Fourier/grid columns -> (member, grid row).  Columns of
// level fields are (field, level, member) with the member fastest; surface
// fields follow as (field, member).
struct ColMap {
  int E, L, nF3, nF2;  // members, levels, level fields, surface fields
  __host__ __device__ int ncols() const { return nF3 * E * L + nF2 * E; }
  __host__ __device__ int rows() const { return nF3 * L + nF2; }
  __host__ __device__ void map(int col, int &e, int &r) const {
    const int EL = E * L;
    if (col < nF3 * EL) {
      const int f = col / EL, rem = col - f * EL;
      const int l = rem / E;
      e = rem - l * E;
      r = f * L + l;
    } else {
      const int c2 = col - nF3 * EL;
      const int g = c2 / E;
      e = c2 - g * E;
      r = nF3 * L + g;
    }
  }
};

// ---------------------------------------------------------------------------
// Length-128 real FFTs, 8 threads per row (64 = 8 x 8 complex FFT of the
// even/odd packed samples, then the real split).  Each transform is split
// into stages separated by a block barrier: the caller places the
// __syncthreads() between stages, so several rows (and other work) share
// the barriers.  scr is an 8x9 complex scratch per row.
// ---------------------------------------------------------------------------
constexpr int FFT_I = 128;
constexpr int FFT_ROWS = 32;

// Twiddles exp(-2 pi i k / 128), k = 0..127, filled by the host.
template <class T>
struct Tw;
template <>
struct Tw<double> {
  static __device__ inline const cx<double> *get();
};
template <>
struct Tw<float> {
  static __device__ inline const cx<float> *get();
};
__constant__ cx<double> c_tw128_d[FFT_I];
__constant__ cx<float> c_tw128_f[FFT_I];
__device__ inline const cx<double> *Tw<double>::get() { return c_tw128_d; }
__device__ inline const cx<float> *Tw<float>::get() { return c_tw128_f; }

template <class T>
using Scr = cx<T> (*)[9];  // [8][9]

// Inverse, stage a: X_m (m < M, from F with stride ms between m) -> scr.
// g = X0.re + 2 sum_m Re(X_m e^{i m lam}) = irfft(zero-filled * nlon)
// (spectral.py:166-172); the imaginary part of X0 is discarded.
template <class T>
__device__ inline void ifft_a(const cx<T> *__restrict__ F, size_t ms, int M, cx<T> (*scr)[9], int tt) {
  const cx<T> *tw = Tw<T>::get();
  cx<T> v[8];
#pragma unroll
  for (int n2 = 0; n2 < 8; ++n2) {
    const int kk = tt + 8 * n2;
    cx<T> A = zero<T>(), B = zero<T>();
    if (kk < M) {
      A = F[(size_t)kk * ms];
      if (kk == 0) A.y = 0;
    }
    const int kb = 64 - kk;
    if (kb < M) B = conj(F[(size_t)kb * ms]);
    const cx<T> w = conj(tw[kk]);  // exp(+2 pi i kk/128)
    v[n2] = (A + B) + cmul(mk((T)0, (T)1), cmul(w, A - B));
  }
  dft8<+1>(v);
#pragma unroll
  for (int k2 = 0; k2 < 8; ++k2) scr[tt][k2] = cmul(v[k2], conj(tw[(2 * tt * k2) & 127]));
}

// Inverse, stage b: scr -> 128 real samples (as 64 complex pairs) in out.
template <class T>
__device__ inline void ifft_b(cx<T> (*scr)[9], cx<T> *out, int tt) {
  cx<T> v[8];
#pragma unroll
  for (int n1 = 0; n1 < 8; ++n1) v[n1] = scr[n1][tt];
  dft8<+1>(v);
#pragma unroll
  for (int k1 = 0; k1 < 8; ++k1) out[tt + 8 * k1] = v[k1];
}

// Forward, stage a: 128 real samples (64 complex pairs in) -> scr.
template <class T>
__device__ inline void fft_a(const cx<T> *in, cx<T> (*scr)[9], int tt) {
  const cx<T> *tw = Tw<T>::get();
  cx<T> v[8];
#pragma unroll
  for (int n2 = 0; n2 < 8; ++n2) v[n2] = in[tt + 8 * n2];
  dft8<-1>(v);
#pragma unroll
  for (int k2 = 0; k2 < 8; ++k2) scr[tt][k2] = cmul(v[k2], tw[(2 * tt * k2) & 127]);
}

// Forward, stage b: scr -> Z[0..63] (complex FFT of the packed samples).
template <class T>
__device__ inline void fft_b(cx<T> (*scr)[9], cx<T> *Z, int tt) {
  cx<T> v[8];
#pragma unroll
  for (int n1 = 0; n1 < 8; ++n1) v[n1] = scr[n1][tt];
  dft8<-1>(v);
#pragma unroll
  for (int k1 = 0; k1 < 8; ++k1) Z[tt + 8 * k1] = v[k1];
}

// Forward, stage c: real split, X_m = rfft/nlon for m < M (spectral.py:159-164),
// written to F with stride ms between m.
template <class T>
__device__ inline void fft_c(const cx<T> *Z, cx<T> *__restrict__ F, size_t ms, int M, int tt) {
  const cx<T> *tw = Tw<T>::get();
  const T half = (T)0.5, inv_n = (T)1 / (T)FFT_I;
#pragma unroll
  for (int k1 = 0; k1 < 8; ++k1) {
    const int k = tt + 8 * k1;
    if (k >= M) break;
    const cx<T> zk = Z[k], zc = conj(Z[(64 - k) & 63]);
    const cx<T> s = zk + zc, d = zk - zc;
    // X = (s - i w d)/2, w = exp(-2 pi i k/128)
    const cx<T> wd = cmul(tw[k], d);
    F[(size_t)k * ms] = inv_n * mk(half * (s.x + wd.y), half * (s.y - wd.x));
  }
}

// Fourier (m <= M-1) -> grid, 32 rows per block; rows are j*C + col.
template <class T>
__global__ __launch_bounds__(256) void k_ifft128(const cx<T> *__restrict__ F, T *__restrict__ G,
                                                 int C, ColMap cm, int J, int M) {
  __shared__ cx<T> sh[FFT_ROWS][8][9];
  const int t = threadIdx.x, rl = t >> 3, tt = t & 7;
  const long row = (long)blockIdx.x * FFT_ROWS + rl;
  const bool live = row < (long)J * C;
  int j = 0, col = 0;
  if (live) {
    j = (int)(row / C);
    col = (int)(row - (long)j * C);
    ifft_a<T>(F + (size_t)j * C + col, (size_t)J * C, M, sh[rl], tt);
  }
  __syncthreads();
  if (!live) return;
  int e, r;
  cm.map(col, e, r);
  ifft_b<T>(sh[rl], reinterpret_cast<cx<T> *>(G + ((size_t)(e * J + j) * cm.rows() + r) * FFT_I), tt);
}

// grid -> Fourier (m <= M-1), 32 rows per block.
template <class T>
__global__ __launch_bounds__(256) void k_fft128(const T *__restrict__ G, cx<T> *__restrict__ F,
                                                int C, ColMap cm, int J, int M) {
  __shared__ cx<T> sh[FFT_ROWS][8][9];
  __shared__ cx<T> zb[FFT_ROWS][64];
  const int t = threadIdx.x, rl = t >> 3, tt = t & 7;
  const long row = (long)blockIdx.x * FFT_ROWS + rl;
  const bool live = row < (long)J * C;
  int j = 0, col = 0;
  if (live) {
    j = (int)(row / C);
    col = (int)(row - (long)j * C);
    int e, r;
    cm.map(col, e, r);
    fft_a<T>(reinterpret_cast<const cx<T> *>(G + ((size_t)(e * J + j) * cm.rows() + r) * FFT_I),
             sh[rl], tt);
  }
  __syncthreads();
  if (live) fft_b<T>(sh[rl], zb[rl], tt);
  __syncthreads();
  if (live) fft_c<T>(zb[rl], F + (size_t)j * C + col, (size_t)J * C, M, tt);
}

// ---------------------------------------------------------------------------
// Legendre synthesis recipes: for each output o, the Fourier value at
// latitude j is  sum_n P_n^m(mu_j) A_o(n) + H_n^m(mu_j) B_o(n), optionally
// times 1/(a cos lat_j).  A_o, B_o are linear in the NIN input fields.
// ---------------------------------------------------------------------------
struct SynLevel {  // (vor, div, t, q) -> (U, V, vor, div, t, q): state.py:94-99, dynamics.py:167-169
  static constexpr int NIN = 4, NOUT = 6;
  static constexpr bool hasA(int) { return true; }
  static constexpr bool hasB(int o) { return o < 2; }
  static constexpr bool scaled(int o) { return o < 2; }
  template <class T>
  __device__ __forceinline__ static void terms(const cx<T> *x, T il, T fm, cx<T> *A, cx<T> *B) {
    const cx<T> psi = il * x[0], chi = il * x[1];
    A[0] = im_mul(fm, chi);  // U = P (im chi) - H psi   (spectral.py:271)
    B[0] = -psi;
    A[1] = im_mul(fm, psi);  // V = P (im psi) + H chi   (spectral.py:272)
    B[1] = chi;
    A[2] = x[0];
    A[3] = x[1];
    A[4] = x[2];
    A[5] = x[3];
  }
};
struct SynSurf {  // (pi, sum dsig vor, sum dsig div) -> (d/dx pi, d/dy pi, pi, sum u, sum v, sum D)
  // hgrad and wave_to_grid of spectral.py:193-225; the column sums of
  // u, v, D (psichi_to_uv and wave_to_grid of the vertically summed spectra)
  static constexpr int NIN = 3, NOUT = 6;
  static constexpr bool hasA(int o) { return o != 1; }
  static constexpr bool hasB(int o) { return o == 1 || o == 3 || o == 4; }
  static constexpr bool scaled(int o) { return o < 2 || o == 3 || o == 4; }
  template <class T>
  __device__ __forceinline__ static void terms(const cx<T> *x, T il, T fm, cx<T> *A, cx<T> *B) {
    A[0] = im_mul(fm, x[0]);
    B[1] = x[0];
    A[2] = x[0];
    const cx<T> psi = il * x[1], chi = il * x[2];
    A[3] = im_mul(fm, chi);
    B[3] = -psi;
    A[4] = im_mul(fm, psi);
    B[4] = chi;
    A[5] = x[2];
  }
};
struct SynScalar {  // wave_to_grid (spectral.py:193-197)
  static constexpr int NIN = 1, NOUT = 1;
  static constexpr bool hasA(int) { return true; }
  static constexpr bool hasB(int) { return false; }
  static constexpr bool scaled(int) { return false; }
  template <class T>
  __device__ __forceinline__ static void terms(const cx<T> *x, T, T, cx<T> *A, cx<T> *) { A[0] = x[0]; }
};
struct SynPsiChi {  // psichi_to_uv (spectral.py:262-276)
  static constexpr int NIN = 2, NOUT = 2;
  static constexpr bool hasA(int) { return true; }
  static constexpr bool hasB(int) { return true; }
  static constexpr bool scaled(int) { return true; }
  template <class T>
  __device__ __forceinline__ static void terms(const cx<T> *x, T, T fm, cx<T> *A, cx<T> *B) {
    A[0] = im_mul(fm, x[1]);
    B[0] = -x[0];
    A[1] = im_mul(fm, x[0]);
    B[1] = x[1];
  }
};
struct SynGrad {  // hgrad (spectral.py:210-225)
  static constexpr int NIN = 1, NOUT = 2;
  static constexpr bool hasA(int o) { return o == 0; }
  static constexpr bool hasB(int o) { return o == 1; }
  static constexpr bool scaled(int) { return true; }
  template <class T>
  __device__ __forceinline__ static void terms(const cx<T> *x, T, T fm, cx<T> *A, cx<T> *B) {
    A[0] = im_mul(fm, x[0]);
    B[1] = x[0];
  }
};

struct OutSlots {
  int off[12];  // column offset in the Fourier row of each output field
};

// ---------------------------------------------------------------------------
// Legendre analysis recipes.  For each input field f the kernel forms
// AP_f = sum_j gw_j P_n^m(mu_j) F_f(j) and (where needed)
// AH_f = sum_j gw_j H_n^m(mu_j) F_f(j); combine() maps them to outputs.
// Vector (u-like) inputs are pre-scaled in grid space by 1/(a cos lat), so
// gw/(a(1-mu^2)) * (u cos lat) of spectral.py:236-243 becomes gw * u/(a cos).
// ---------------------------------------------------------------------------
struct AnaTend {  // (FU, FV, KE, XU, XV, GT, YU, YV, G) -> (vor, div, t, q): dynamics.py:219,246-262
  static constexpr int NIN = 9, NOUT = 4;
  static constexpr bool ACC = false;
  static constexpr bool needP(int f) { return f != 4 && f != 7; }
  static constexpr bool needH(int f) { return f == 0 || f == 1 || f == 4 || f == 7; }
  template <class T>
  __device__ __forceinline__ static void combine(const cx<T> *AP, const cx<T> *AH, T fm, T nn1a2, cx<T> *o) {
    o[0] = im_mul(fm, AP[1]) + AH[0];                                   // curl(fu, fv)
    o[1] = im_mul(fm, AP[0]) - AH[1] + nn1a2 * AP[2];                   // div - lap(ke)
    o[2] = AP[5] - im_mul(fm, AP[3]) + AH[4];                           // -div(xu,xv) + gt
    o[3] = AP[8] - im_mul(fm, AP[6]) + AH[7];                           // -div(yu,yv) + g
  }
};
struct AnaState {  // (u', v', t, q) -> (vor, div, t, q): state.py:86-92
  static constexpr int NIN = 4, NOUT = 4;
  static constexpr bool ACC = false;
  static constexpr bool needP(int) { return true; }
  static constexpr bool needH(int f) { return f < 2; }
  template <class T>
  __device__ __forceinline__ static void combine(const cx<T> *AP, const cx<T> *AH, T fm, T, cx<T> *o) {
    o[0] = im_mul(fm, AP[1]) + AH[0];
    o[1] = im_mul(fm, AP[0]) - AH[1];
    o[2] = AP[2];
    o[3] = AP[3];
  }
};
struct AnaVorDiv {  // uv_to_vordiv (spectral.py:227-244)
  static constexpr int NIN = 2, NOUT = 2;
  static constexpr bool ACC = false;
  static constexpr bool needP(int) { return true; }
  static constexpr bool needH(int) { return true; }
  template <class T>
  __device__ __forceinline__ static void combine(const cx<T> *AP, const cx<T> *AH, T fm, T, cx<T> *o) {
    o[0] = im_mul(fm, AP[1]) + AH[0];
    o[1] = im_mul(fm, AP[0]) - AH[1];
  }
};
struct AnaScalar {  // grid_to_wave (spectral.py:186-191)
  static constexpr int NIN = 1, NOUT = 1;
  static constexpr bool ACC = false;
  static constexpr bool needP(int) { return true; }
  static constexpr bool needH(int) { return false; }
  template <class T>
  __device__ __forceinline__ static void combine(const cx<T> *AP, const cx<T> *, T, T, cx<T> *o) { o[0] = AP[0]; }
};

struct InSlots {
  int off[12];  // column offset in the Fourier row of each input field
};

}  // namespace lbm
