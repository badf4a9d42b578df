// Device code of the gfx950 time-stepping core: complex helpers, the
// real FFT pair (128 and 384 longitudes specialised, other lengths of
// radices 2, 3, 4, 5 through a Stockham pass loop), the per-m Legendre synthesis/analysis with
// north/south folding, the fused grid-point tendency kernel and the fused
// spectral semi-implicit leapfrog + vertical diffusion + Robert-Asselin
// kernel.  Everything is templated on the arithmetic type T (double/float).
//
// Internal layouts (E members, L levels, M = nmax+1, J lats, Jh = J/2):
//   spectral  [s][f][ncols]  complex, s = packed (m, n>=m) index,
//             ncols = E*L for level fields, E for surface fields
//   Fourier   [m][C/8][j][8] complex, C = nF3*E*L + nF2*E columns padded
//             to a multiple of 8, in tiles of 8 columns (fidx)
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

// Streaming loads of data read once per step (written by the previous
// kernel, read by this one only).
template <class T>
__device__ __forceinline__ T ld_nt(const T *p) {
  return __builtin_nontemporal_load(p);
}
template <class T>
__device__ __forceinline__ cx<T> ld_nt(const cx<T> *p) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  const v2 w = __builtin_nontemporal_load(reinterpret_cast<const v2 *>(p));
  return mk(w.x, w.y);
}

// Raw buffer view of complex values: the resource (base, byte range) sits in
// scalar registers, each access adds a 32-bit per-lane byte offset and a
// wave-uniform byte offset (soffset).  A kernel that walks one array at a
// runtime stride (the level stride of the spectral state) then needs one
// offset register per lane instead of a 64-bit address per level.  Accesses
// at or past the range are dropped by the hardware (loads return zero).
struct BufView {
  __amdgpu_buffer_rsrc_t r;
  // 0x00020000: the raw (untyped, 32-bit data format) buffer word of gfx9
  __device__ __forceinline__ BufView(const void *p, unsigned bytes)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000)) {}
  __device__ __forceinline__ cx<double> ld(unsigned voff, unsigned soff, const cx<double> *) const {
    typedef int i4 __attribute__((ext_vector_type(4)));
    const i4 w = __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0);
    return __builtin_bit_cast(cx<double>, w);
  }
  __device__ __forceinline__ cx<float> ld(unsigned voff, unsigned soff, const cx<float> *) const {
    typedef int i2 __attribute__((ext_vector_type(2)));
    const i2 w = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0);
    return __builtin_bit_cast(cx<float>, w);
  }
  // non-temporal loads (aux bit 1: the data is not expected to be reused)
  __device__ __forceinline__ cx<double> ld_nt(unsigned voff, unsigned soff, const cx<double> *) const {
    typedef int i4 __attribute__((ext_vector_type(4)));
    const i4 w = __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 2);
    return __builtin_bit_cast(cx<double>, w);
  }
  __device__ __forceinline__ cx<float> ld_nt(unsigned voff, unsigned soff, const cx<float> *) const {
    typedef int i2 __attribute__((ext_vector_type(2)));
    const i2 w = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 2);
    return __builtin_bit_cast(cx<float>, w);
  }
  // real values; aux 2 = non-temporal
  template <int AUX = 0>
  __device__ __forceinline__ double ld(unsigned voff, unsigned soff, const double *) const {
    typedef int i2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, AUX));
  }
  template <int AUX = 0>
  __device__ __forceinline__ float ld(unsigned voff, unsigned soff, const float *) const {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, AUX));
  }
  template <int AUX = 0>
  __device__ __forceinline__ void st(unsigned voff, unsigned soff, double v) const {
    typedef int i2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(i2, v), r, (int)voff, (int)soff, AUX);
  }
  template <int AUX = 0>
  __device__ __forceinline__ void st(unsigned voff, unsigned soff, float v) const {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, (int)voff, (int)soff, AUX);
  }
  __device__ __forceinline__ void st(unsigned voff, unsigned soff, cx<double> v) const {
    typedef int i4 __attribute__((ext_vector_type(4)));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(i4, v), r, (int)voff, (int)soff, 0);
  }
  __device__ __forceinline__ void st(unsigned voff, unsigned soff, cx<float> v) const {
    typedef int i2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(i2, v), r, (int)voff, (int)soff, 0);
  }
};

// Streaming stores of data read back only by the next kernel (from HBM):
// non-temporal, so they do not displace reused lines from the caches.
template <class T>
__device__ __forceinline__ void st_nt(T *p, T v) {
  __builtin_nontemporal_store(v, p);
}
template <class T>
__device__ __forceinline__ void st_nt(cx<T> *p, cx<T> v) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  v2 w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<v2 *>(p));
}
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
// Fourier arrays are [m][column tile of 8][j][8] complex, the C columns
// padded to a multiple of 8: one m of a tile is J x 8 contiguous values
// (what a Legendre block of 8 columns reads or writes), and 8 columns x 4
// latitudes of an FFT block are 32 contiguous values.
__host__ __device__ __forceinline__ int fpad(int C) { return (C + 7) & ~7; }
__host__ __device__ __forceinline__ size_t fidx(int m, int j, int c, int J, int C) {
  return (((size_t)m * (fpad(C) >> 3) + (c >> 3)) * J + j) * 8 + (c & 7);
}

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

// Members still advancing at the current step of a run: member e is live
// while *step < inistp + nsteps[e], or while a deferred surface-pressure
// update of its last step is pending (pend[e] != 0, applied by the next
// tendency call from that step's synthesised pi); all members when step is
// null.  The time-loop kernels skip the work of finished members (whole
// blocks when all their members have finished); read on the device so a
// captured graph sees the step of each replay.
//
// ecnt (a multiple of 8 below E, or 0 for all) limits a launch to the first
// ecnt members: the host knows the lengths and orders nothing, so the
// members that can still be live must be a prefix for this to pay; columns
// (field-level, member) then map to blocks as (field-level, 8-member tile).
struct ActiveMask {
  const int *step;    // the loop counter, or null
  const int *nsteps;  // [E] full steps per member
  const int *pend;    // [E] pending ps' updates, or null
  int inistp, E;
  int ecnt;           // members covered by the launch (0: all)
  __host__ __device__ __forceinline__ bool reduced() const { return ecnt > 0 && ecnt < E; }
  __device__ __forceinline__ bool operator()(int e) const {
    return step == nullptr || *step < inistp + nsteps[e] || (pend != nullptr && pend[e] != 0);
  }
  // first column of 8-column tile bx; columns are (field-level) * E + member
  __device__ __forceinline__ int tile_col0(int bx) const {
    if (!reduced()) return bx * 8;
    const int tpl = ecnt >> 3;
    const int fl = bx / tpl;
    return fl * E + (bx - fl * tpl) * 8;
  }
  // row of a transform over J latitudes and C columns -> (j, col); false
  // past the last row of the launch and for the padding columns.  Rows run
  // over 8 columns, then latitudes, then column tiles (the Fourier layout);
  // a reduced launch covers the tiles of the first ecnt members of each
  // field-level (E is then a multiple of 8).
  __device__ __forceinline__ bool row_col(long row, int J, int C, int &j, int &col) const {
    const int c8 = (int)(row & 7);
    const long r8 = row >> 3;
    if (!reduced()) {
      if (r8 >= (long)(fpad(C) >> 3) * J) return false;
      const int tile = (int)(r8 / J);
      j = (int)(r8 - (long)tile * J);
      col = tile * 8 + c8;
      return col < C;
    }
    const int nfl = C / E, tpl = ecnt >> 3;
    if (r8 >= (long)nfl * tpl * J) return false;
    const int t = (int)(r8 / J);
    j = (int)(r8 - (long)t * J);
    const int fl = t / tpl;
    col = fl * E + (t - fl * tpl) * 8 + c8;
    return true;
  }
  // any member among columns col0 .. col0+7 (< ncols), member = col % E
  __device__ __forceinline__ bool any8(int col0, int ncols) const {
    if (step == nullptr) return true;
    bool any = false;
#pragma unroll
    for (int c = 0; c < 8; ++c)
      if (col0 + c < ncols && (*this)((col0 + c) % E)) any = true;
    return any;
  }
};

// ---------------------------------------------------------------------------
// Real FFTs of length NI = 128 R (R = 1: 128 longitudes, R = 3: 384).  The
// NI real samples of a row are NC = NI/2 complex pairs z_n = x_2n + i x_2n+1;
// their complex DFT of length NC = 64 R is split as
//   n = n_b + 64 n_a,  k = r + R q   (n_b, q < 64; n_a, r < R)
// into R 64-point DFTs (over q for the inverse, over n_b for the forward),
// twiddles exp(-+2 pi i n_b r / NC) and a radix-R DFT across r / n_a.  Each
// 64-point DFT is 8 x 8 (radix 8, LDS transpose) on 8 threads, so a row
// uses 8R threads; then the real split to/from X_m = rfft/nlon.  Stages are
// separated by block barriers placed by the kernels.
// ---------------------------------------------------------------------------
template <class T, int NI, bool INV = false, int RW = 0>
struct FftShape {
  static constexpr int R = NI / 128;
  static constexpr int NC = NI / 2;
  static constexpr int TPR = 8 * R;             // threads per row
  // rows per block (INV: the inverse transform): consecutive rows are
  // consecutive Fourier columns, so a block's accesses to one coefficient m
  // form runs of 8 x sizeof(cx<T>) bytes; the LDS per row (1.2 KB FP64 at
  // 128 points, 3.4 KB at 384) bounds the blocks per CU.  Measured best:
  // FP64 16 rows at 128 points (fewer waves per barrier at the same
  // occupancy), 10 (inverse) and 20 (forward) at 384; FP32 32 rows.
  // (RW > 0: an autotuner variant with RW rows per block)
  static constexpr int ROWS = RW > 0 ? RW : sizeof(T) == 4 ? 32 : (R == 1 ? 16 : (INV ? 10 : 20));
  static constexpr int THREADS = ROWS * TPR;
  static constexpr int LDS = 8 * 9 * R;         // complex scratch per row (>= NC)
};

// Twiddles exp(-2 pi i k / NI), k = 0..NI-1, filled by the host.
// (static: each translation unit that includes this header has its own)
// The stage functions below take the table (this one or a copy in LDS) as
// tw.
static __constant__ cx<double> c_tw128_d[128];
static __constant__ cx<float> c_tw128_f[128];
static __constant__ cx<double> c_tw384_d[384];
static __constant__ cx<float> c_tw384_f[384];
template <class T, int NI>
__device__ inline const cx<T> *twiddles();
template <>
__device__ inline const cx<double> *twiddles<double, 128>() { return c_tw128_d; }
template <>
__device__ inline const cx<float> *twiddles<float, 128>() { return c_tw128_f; }
template <>
__device__ inline const cx<double> *twiddles<double, 384>() { return c_tw384_d; }
template <>
__device__ inline const cx<float> *twiddles<float, 384>() { return c_tw384_f; }

// exp(SIGN 2 pi i a / b) from the table, b dividing NI
template <class T, int NI, int SIGN>
__device__ __forceinline__ cx<T> twf(const cx<T> *tw, int a, int b) {
  const cx<T> w = tw[(a * (NI / b)) % NI];
  return SIGN < 0 ? w : conj(w);
}

// radix-R DFT of v[0..R-1] in place (R = 1 or 3), sign SIGN
template <int R, int SIGN, class T>
__device__ __forceinline__ void dftR(cx<T> *v) {
  if constexpr (R == 3) {
    // exp(SIGN 2 pi i / 3) = -1/2 + SIGN i sqrt(3)/2
    const T c = (T)-0.5, sn = (T)SIGN * (T)0.86602540378443864676;
    const cx<T> s = v[1] + v[2], d = v[1] - v[2];
    const cx<T> m = v[0] + c * s;
    const cx<T> jd = mk(-sn * d.y, sn * d.x);  // i sn d
    v[0] = v[0] + s;
    v[1] = m + jd;
    v[2] = m - jd;
  }
}

// Scale of the Fourier coefficient m applied by the transforms: mode 0
// none (the forward keeps its 1/nlon); 1 the transpose of rfft/nlon seen
// through the real inverse ((1 or 1/2)/nlon, on the inverse's input); 2 the
// transpose of the real inverse (1 or 2, on the forward's unnormalised
// output in place of 1/nlon).
template <class T, int NI>
__device__ __forceinline__ T fourier_scale(int mode, int m) {
  if (mode == 1) return (m == 0 ? (T)1 : (T)0.5) / (T)NI;
  if (mode == 2) return m == 0 ? (T)1 : (T)2;
  return (T)1;
}

// First-pass outputs v[k2] of thread (r, t) into the row's scratch: the
// 8 x 8 transpose of sub-transform r, rows padded to 9 (distinct banks for
// the 8 lanes of a row in both directions).
template <class T, int NI>
__device__ inline void fft_a_store(const cx<T> *v, cx<T> *scr, int tt) {
  const int r = tt >> 3, t = tt & 7;
  cx<T> *sc = scr + r * 72 + t * 9;
#pragma unroll
  for (int k2 = 0; k2 < 8; ++k2) sc[k2] = v[k2];
}

// Inverse, stage a (thread (r, t) of the row): X_m = X(m) (m < M) -> Z_k
// for k = r + R(t + 8 n2) -> first radix-8 pass, twiddled, in v (to scr by
// fft_a_store).
// g = X0.re + 2 sum_m Re(X_m e^{i m lam}) = irfft(zero-filled * nlon)
// (spectral.py:166-172); the imaginary part of X0 is discarded.
template <class T, int NI, class XF>
__device__ inline void ifft_a(XF X, int M, cx<T> *v, int tt, int mode, const cx<T> *tw) {
  using S = FftShape<T, NI>;
  const int r = tt >> 3, t = tt & 7;
#pragma unroll
  for (int n2 = 0; n2 < 8; ++n2) {
    const int k = r + S::R * (t + 8 * n2);
    cx<T> A = zero<T>(), B = zero<T>();
    if (k < M) {
      A = fourier_scale<T, NI>(mode, k) * X(k);
      if (k == 0) A.y = 0;
    }
    const int kb = S::NC - k;
    if (kb < M) B = conj(fourier_scale<T, NI>(mode, kb) * X(kb));
    const cx<T> w = conj(tw[k]);  // exp(+2 pi i k / NI)
    v[n2] = (A + B) + cmul(mk((T)0, (T)1), cmul(w, A - B));
  }
  dft8<+1>(v);
#pragma unroll
  for (int k2 = 0; k2 < 8; ++k2) v[k2] = cmul(v[k2], twf<T, NI, +1>(tw, t * k2, 64));
}

// Inverse, stage b: second radix-8 pass of sub-transform r, then the
// twiddle exp(+2 pi i n_b r / NC); returns Y_r[n_b = t + 8 k1] in v.
template <class T, int NI>
__device__ inline void ifft_b(const cx<T> *scr, cx<T> *v, int tt, const cx<T> *tw) {
  const int r = tt >> 3, t = tt & 7;
#pragma unroll
  for (int n1 = 0; n1 < 8; ++n1) v[n1] = scr[r * 72 + n1 * 9 + t];
  dft8<+1>(v);
  if (FftShape<T, NI>::R > 1) {
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) v[k1] = cmul(v[k1], twf<T, NI, +1>(tw, (t + 8 * k1) * r, FftShape<T, NI>::NC));
  }
}

// Inverse, stage c (R > 1): radix-R across r from ex[r][n_b] for output
// block n_a = r of this thread: o[k1] = z[n_b + 64 n_a], n_b = t + 8 k1.
template <class T, int NI>
__device__ inline void ifft_c(const cx<T> *ex, cx<T> *o, int tt) {
  constexpr int R = FftShape<T, NI>::R;
  const int na = tt >> 3, t = tt & 7;
#pragma unroll
  for (int k1 = 0; k1 < 8; ++k1) {
    const int nb = t + 8 * k1;
    cx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = ex[r * 64 + nb];
    dftR<R, +1>(v);
    o[k1] = v[0];
#pragma unroll
    for (int q = 1; q < R; ++q)
      if (q == na) o[k1] = v[q];
  }
}

// A row of NC complex samples from LDS to memory, two per lane with
// non-temporal 16-byte stores (FP32: full 128-byte lines).
template <class T, int NI>
__device__ inline void store_row_pairs(const cx<T> *z, cx<T> *out, int tt) {
  using S = FftShape<T, NI>;
  typedef T t4 __attribute__((ext_vector_type(4)));
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int i = tt + S::TPR * q;  // samples 2i, 2i + 1 (NC / 2 = 4 TPR pairs)
    const cx<T> a0 = z[2 * i], a1 = z[2 * i + 1];
    t4 w;
    w.x = a0.x;
    w.y = a0.y;
    w.z = a1.x;
    w.w = a1.y;
    __builtin_nontemporal_store(w, reinterpret_cast<t4 *>(out + 2 * i));
  }
}

// Forward, stage a (thread (r, t)): radix-R across n_a of z[n_b + 64 n_a]
// for n_b = t + 8 n2, keep output r, twiddle exp(-2 pi i n_b r / NC), first
// radix-8 pass; v[k2] (twiddled) goes to scr by fft_a_store.  in: the row in
// global memory (R = 1) or staged in LDS (R > 1).
template <class T, int NI, bool LDS_IN = false>
__device__ inline void fft_a(const cx<T> *in, cx<T> *v, int tt, const cx<T> *tw) {
  constexpr int R = FftShape<T, NI>::R;
  const int r = tt >> 3, t = tt & 7;
#pragma unroll
  for (int n2 = 0; n2 < 8; ++n2) {
    const int nb = t + 8 * n2;
    if constexpr (R == 1) {
      // FP64 rows stream past the caches (measured: FP32 rows lose); a row
      // staged in LDS (LDS_IN) is read directly
      if constexpr (sizeof(T) == 8 && !LDS_IN)
        v[n2] = ld_nt(in + nb);
      else
        v[n2] = in[nb];
    } else {
      cx<T> u[R];
#pragma unroll
      for (int na = 0; na < R; ++na) u[na] = in[nb + 64 * na];
      dftR<R, -1>(u);
      cx<T> ur = u[0];
#pragma unroll
      for (int q = 1; q < R; ++q)
        if (q == r) ur = u[q];
      v[n2] = cmul(ur, twf<T, NI, -1>(tw, nb * r, FftShape<T, NI>::NC));
    }
  }
  dft8<-1>(v);
#pragma unroll
  for (int k2 = 0; k2 < 8; ++k2) v[k2] = cmul(v[k2], twf<T, NI, -1>(tw, t * k2, 64));
}
// Forward, stage b: second radix-8 pass; v[k1] = Z[r + R (t + 8 k1)].
template <class T, int NI>
__device__ inline void fft_b(const cx<T> *scr, cx<T> *v, int tt) {
  const int r = tt >> 3, t = tt & 7;
#pragma unroll
  for (int n1 = 0; n1 < 8; ++n1) v[n1] = scr[r * 72 + n1 * 9 + t];
  dft8<-1>(v);
}

// Forward, stage c: real split, X_m = rfft/nlon for m < M (spectral.py:159-164)
// (times the mode's scale), written to F with stride ms between m.
template <class T, int NI>
__device__ inline void fft_c(const cx<T> *Z, cx<T> *__restrict__ F, size_t ms, int M, int tt,
                             int mode, const cx<T> *tw) {
  using S = FftShape<T, NI>;
  const T half = (T)0.5;
  for (int k = tt; k < M; k += S::TPR) {
    const cx<T> zk = Z[k], zc = conj(Z[(S::NC - k) % S::NC]);
    const cx<T> s = zk + zc, d = zk - zc;
    // X = (s - i w d)/2, w = exp(-2 pi i k / NI)
    const cx<T> wd = cmul(tw[k], d);
    const T sc = mode == 2 ? fourier_scale<T, NI>(2, k) : (T)1 / (T)NI;
    F[(size_t)k * ms] = sc * mk(half * (s.x + wd.y), half * (s.y - wd.x));
  }
}

// Fourier (m <= M-1) -> grid; rows are j*C + col.  mode: fourier_scale.
template <class T, int NI, int RW = 0>
__global__ __launch_bounds__((FftShape<T, NI, true, RW>::THREADS)) void k_ifft(const cx<T> *__restrict__ F,
                                                                T *__restrict__ G, int C, ColMap cm,
                                                                int J, int M, int mode,
                                                                ActiveMask am) {
  using S = FftShape<T, NI, true, RW>;
  __shared__ cx<T> sh[S::ROWS][S::LDS];
  const int t = threadIdx.x, rl = t / S::TPR, tt = t - rl * S::TPR;
  const long row = (long)blockIdx.x * S::ROWS + rl;
  int j = 0, col = 0, e = 0, r = 0;
  bool live = am.row_col(row, J, C, j, col);
  if (live) {
    cm.map(col, e, r);
    live = am(e);
  }
  // The row's M coefficients staged in LDS once (ifft_a reads X_m and
  // X_{NC-m} of each m); M <= NC (nmax + 1 <= nlon / 2, checked by the
  // plan).  384 FP32 points read them from memory (measured: staging loses).
  constexpr bool STAGE = S::R == 1 || sizeof(T) == 8;
  static_assert(S::LDS >= S::NC, "the row scratch holds the staged coefficients");
  const cx<T> *Fr = F + fidx(0, j, col, J, C);
  const size_t ms = (size_t)J * fpad(C);
  cx<T> v[8];
  if constexpr (STAGE) {
    if (live)
      for (int m = tt; m < M; m += S::TPR) sh[rl][m] = Fr[(size_t)m * ms];
    __syncthreads();
    const cx<T> *X = sh[rl];
    if (live) ifft_a<T, NI>([&](int k) { return X[k]; }, M, v, tt, mode, twiddles<T, NI>());
    __syncthreads();  // the staged coefficients are overwritten by the first pass
  } else {
    if (live) ifft_a<T, NI>([&](int k) { return Fr[(size_t)k * ms]; }, M, v, tt, mode, twiddles<T, NI>());
  }
  if (live) fft_a_store<T, NI>(v, sh[rl], tt);
  __syncthreads();
  if (live) ifft_b<T, NI>(sh[rl], v, tt, twiddles<T, NI>());
  cx<T> *out = reinterpret_cast<cx<T> *>(G + ((size_t)(e * J + j) * cm.rows() + r) * NI);
  if constexpr (S::R == 1 && sizeof(T) == 8) {
    if (live) {
#pragma unroll
      for (int k1 = 0; k1 < 8; ++k1) st_nt(out + tt + 8 * k1, v[k1]);
    }
  } else if constexpr (S::R == 1) {
    // FP32: the row back through LDS in natural order, so that each lane
    // stores two consecutive values (full 128-byte lines per row and store)
    __syncthreads();  // scr is reused
    if (live) {
#pragma unroll
      for (int k1 = 0; k1 < 8; ++k1) sh[rl][tt + 8 * k1] = v[k1];
    }
    __syncthreads();
    if (live) store_row_pairs<T, NI>(sh[rl], out, tt);
  } else {
    __syncthreads();  // scr is reused as the exchange buffer
    if (live) {
#pragma unroll
      for (int k1 = 0; k1 < 8; ++k1) sh[rl][(tt >> 3) * 64 + (tt & 7) + 8 * k1] = v[k1];
    }
    __syncthreads();
    cx<T> o[8];
    if (live) ifft_c<T, NI>(sh[rl], o, tt);
    const int na = tt >> 3, t8 = tt & 7;
    if constexpr (sizeof(T) == 8) {
      if (live) {
#pragma unroll
        for (int k1 = 0; k1 < 8; ++k1) st_nt(out + t8 + 8 * k1 + 64 * na, o[k1]);
      }
    } else {
      __syncthreads();  // the exchange buffer takes the row in natural order
      if (live) {
#pragma unroll
        for (int k1 = 0; k1 < 8; ++k1) sh[rl][t8 + 8 * k1 + 64 * na] = o[k1];
      }
      __syncthreads();
      if (live) store_row_pairs<T, NI>(sh[rl], out, tt);
    }
  }
}

// grid -> Fourier (m <= M-1).  mode: fourier_scale.
template <class T, int NI, int RW = 0>
__global__ __launch_bounds__((FftShape<T, NI, false, RW>::THREADS)) void k_fft(const T *__restrict__ G,
                                                               cx<T> *__restrict__ F, int C,
                                                               ColMap cm, int J, int M, int mode,
                                                               ActiveMask am) {
  using S = FftShape<T, NI, false, RW>;
  __shared__ cx<T> sh[S::ROWS][S::LDS];
  const int t = threadIdx.x, rl = t / S::TPR, tt = t - rl * S::TPR;
  const long row = (long)blockIdx.x * S::ROWS + rl;
  int j = 0, col = 0, e = 0, r = 0;
  bool live = am.row_col(row, J, C, j, col);
  if (live) {
    cm.map(col, e, r);
    live = am(e);
  }
  const cx<T> *row_in = reinterpret_cast<const cx<T> *>(G + ((size_t)(e * J + j) * cm.rows() + r) * NI);
  cx<T> v[8];
  if constexpr (S::R == 1 && sizeof(T) == 8) {
    if (live) {
      fft_a<T, NI>(row_in, v, tt, twiddles<T, NI>());
      fft_a_store<T, NI>(v, sh[rl], tt);
    }
  } else {
    // the row's 64 R samples staged once in LDS: each of the R threads of a
    // column needs all R of its inputs, and FP32 lanes load two consecutive
    // values (full 128-byte lines)
    if (live) {
      if constexpr (sizeof(T) == 4) {
        typedef T t4 __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int i = tt + S::TPR * q;  // pair of samples 2i, 2i + 1
          const t4 w = *reinterpret_cast<const t4 *>(row_in + 2 * i);
          sh[rl][2 * i] = mk(w.x, w.y);
          sh[rl][2 * i + 1] = mk(w.z, w.w);
        }
      } else {
        const int ra = tt >> 3, ta = tt & 7;
#pragma unroll
        for (int n2 = 0; n2 < 8; ++n2) sh[rl][ta + 8 * n2 + 64 * ra] = row_in[ta + 8 * n2 + 64 * ra];
      }
    }
    __syncthreads();
    if (live) fft_a<T, NI>(sh[rl], v, tt, twiddles<T, NI>());
    __syncthreads();  // the staged row is overwritten by the first pass
    if (live) fft_a_store<T, NI>(v, sh[rl], tt);
  }
  __syncthreads();
  if (live) fft_b<T, NI>(sh[rl], v, tt);
  __syncthreads();  // scr is reused for Z
  if (live) {
    const int rr = tt >> 3, tq = tt & 7;
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) sh[rl][rr + S::R * (tq + 8 * k1)] = v[k1];
  }
  __syncthreads();
  if (live) fft_c<T, NI>(sh[rl], F + fidx(0, j, col, J, C), (size_t)J * fpad(C), M, tt, mode,
                         twiddles<T, NI>());
}

// ---------------------------------------------------------------------------
// Real FFTs of the other grids: NI = 2 NC longitudes with NC a product of
// the radices 2, 3, 4 and 5, NC <= 512 (T21 64, T63 192, T85 256, T106 320
// ...).  Same row addressing, real split and scale modes as k_ifft / k_fft;
// the complex DFT of length NC is a Stockham autosort in LDS, one pass per
// radix, ping-ponging between two buffers of the row: pass q (radix r,
// Ns = product of the earlier radices) maps
//   v_t = a[j + t NC/r] exp(SIGN 2 pi i t (j mod Ns) / (Ns r)),  t < r,
//   b[(j / Ns) Ns r + (j mod Ns) + t Ns] = (radix-r DFT of v)_t
// for j < NC/r, and leaves the transform in natural order.  Twiddles
// exp(-2 pi i k / NI) come from a table in global memory.
// ---------------------------------------------------------------------------
struct FftFactors {
  int NI, NC, np;  // longitudes, NC = NI / 2, passes
  int r[12];       // radix of each pass
  // Ns of each pass (product of the earlier radices) and ceil(2^20 / Ns):
  // j / Ns = (j * dmul) >> 20 exactly for j < 512
  int ns[12], dmul[12];
};
constexpr int FFT_ANY_TPR = 32;  // threads per row

template <int SIGN, class T>
__device__ __forceinline__ void dft5(cx<T> *v) {
  // exp(SIGN 2 pi i k / 5) = c_k + SIGN i s_k
  const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
  const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
  const cx<T> t1 = v[1] + v[4], t2 = v[2] + v[3], t3 = v[1] - v[4], t4 = v[2] - v[3];
  const cx<T> a1 = v[0] + c1 * t1 + c2 * t2, a2 = v[0] + c2 * t1 + c1 * t2;
  const cx<T> b1 = (T)SIGN * (s1 * t3 + s2 * t4), b2 = (T)SIGN * (s2 * t3 - s1 * t4);
  v[0] = v[0] + t1 + t2;
  v[1] = a1 + mk(-b1.y, b1.x);
  v[4] = a1 - mk(-b1.y, b1.x);
  v[2] = a2 + mk(-b2.y, b2.x);
  v[3] = a2 - mk(-b2.y, b2.x);
}

template <int SIGN, class T>
__device__ __forceinline__ void dft_radix(cx<T> *v, int r) {
  if (r == 4) {
    dft4<SIGN>(v[0], v[1], v[2], v[3]);
  } else if (r == 2) {
    const cx<T> a = v[0];
    v[0] = a + v[1];
    v[1] = a - v[1];
  } else if (r == 3) {
    dftR<3, SIGN>(v);
  } else {
    dft5<SIGN>(v);
  }
}

// One row's complex DFT of length ff.NC (sign SIGN, unnormalised) from a
// into LDS buffers a / b; returns the buffer that holds the result.  Every
// thread of the block calls it (the passes end in block barriers).
template <int SIGN, class T>
__device__ cx<T> *stockham_row(cx<T> *a, cx<T> *b, const FftFactors &ff, const cx<T> *__restrict__ tw,
                               int tt, bool live) {
  const int NC = ff.NC, NI = ff.NI;
  for (int q = 0; q < ff.np; ++q) {
    const int r = ff.r[q], Ns = ff.ns[q], nr = NC / r, step = NI / (Ns * r), dm = ff.dmul[q];
    if (live) {
      for (int j = tt; j < nr; j += FFT_ANY_TPR) {
        const int g = (j * dm) >> 20, k = j - g * Ns;
        cx<T> v[5];
        v[0] = a[j];
#pragma unroll
        for (int t = 1; t < 5; ++t) {
          v[t] = zero<T>();
          if (t < r) {
            // k t step < Ns r step = NI: no wrap
            const cx<T> w = tw[k * t * step];
            v[t] = cmul(a[j + t * nr], SIGN < 0 ? w : conj(w));
          }
        }
        dft_radix<SIGN>(v, r);
        const int base = g * Ns * r + k;
#pragma unroll
        for (int t = 0; t < 5; ++t)
          if (t < r) b[base + t * Ns] = v[t];
      }
    }
    __syncthreads();
    cx<T> *c = a;
    a = b;
    b = c;
  }
  return a;
}

template <class T>
__device__ __forceinline__ T fourier_scale_any(int mode, int m, int NI) {
  if (mode == 1) return (m == 0 ? (T)1 : (T)0.5) / (T)NI;
  if (mode == 2) return m == 0 ? (T)1 : (T)2;
  return (T)1;
}

constexpr int FFT_ANY_ROWS = 8;  // rows per block: one tile of 8 Fourier columns

// The rows of a Stockham FFT block: each row's Fourier column base
// (fidx(0, j, col)), -1 when the row is past the launch or its member has
// finished, and its grid row (e, j, r).  Threads 0..7 fill them.
struct AnyRows {
  long fbase[FFT_ANY_ROWS];
  long gbase[FFT_ANY_ROWS];
};
__device__ inline void any_rows(AnyRows &rs, int C, const ColMap &cm, int J, int NI,
                                const ActiveMask &am) {
  if (threadIdx.x < FFT_ANY_ROWS) {
    const int q = threadIdx.x;
    const long row = (long)blockIdx.x * FFT_ANY_ROWS + q;
    int j = 0, col = 0, e = 0, r = 0;
    bool live = am.row_col(row, J, C, j, col);
    if (live) {
      cm.map(col, e, r);
      live = am(e);
    }
    rs.fbase[q] = live ? (long)fidx(0, j, col, J, C) : -1;
    rs.gbase[q] = live ? (long)((size_t)(e * J + j) * cm.rows() + r) * NI : -1;
  }
}

// Fourier (m <= M-1) -> grid for any FftFactors length: 8 rows (one tile of
// Fourier columns) per block, 32 threads and 2 NC complex of dynamic LDS per
// row.  The block stages the 8 columns' coefficients together, m by m, so
// each coefficient's 8 columns are one contiguous run.
template <class T>
__global__ __launch_bounds__(256) void k_ifft_any(const cx<T> *__restrict__ F, T *__restrict__ G, int C,
                                                  ColMap cm, int J, int M, int mode, ActiveMask am,
                                                  FftFactors ff, const cx<T> *__restrict__ tw) {
  extern __shared__ unsigned char fft_any_lds[];
  __shared__ AnyRows rs;
  const int NC = ff.NC, NI = ff.NI;
  const int rl = threadIdx.x / FFT_ANY_TPR, tt = threadIdx.x - rl * FFT_ANY_TPR;
  cx<T> *a = reinterpret_cast<cx<T> *>(fft_any_lds) + (size_t)rl * 2 * NC, *b = a + NC;
  // the twiddle table after the rows
  cx<T> *twl = reinterpret_cast<cx<T> *>(fft_any_lds) + (size_t)FFT_ANY_ROWS * 2 * NC;
  for (int k = threadIdx.x; k < NI; k += blockDim.x) twl[k] = tw[k];
  any_rows(rs, C, cm, J, NI, am);
  __syncthreads();
  tw = twl;
  const size_t ms = (size_t)J * fpad(C);
  for (int idx = threadIdx.x; idx < FFT_ANY_ROWS * M; idx += blockDim.x) {
    const int q = idx & (FFT_ANY_ROWS - 1), m = idx / FFT_ANY_ROWS;
    if (rs.fbase[q] >= 0)
      reinterpret_cast<cx<T> *>(fft_any_lds)[(size_t)q * 2 * NC + NC + m] = F[rs.fbase[q] + (size_t)m * ms];
  }
  __syncthreads();
  const bool live = rs.fbase[rl] >= 0;
  if (live) {
    // Z_k = (A + B) + i e^{+2 pi i k / NI} (A - B), A = X_k, B = conj X_{NC-k}
    // (spectral.py:166-172: g = X0.re + 2 sum_m Re(X_m e^{i m lam}))
    for (int k = tt; k < NC; k += FFT_ANY_TPR) {
      cx<T> A = zero<T>(), B = zero<T>();
      if (k < M) {
        A = fourier_scale_any<T>(mode, k, NI) * b[k];
        if (k == 0) A.y = 0;
      }
      const int kb = NC - k;
      if (kb < M) B = conj(fourier_scale_any<T>(mode, kb, NI) * b[kb]);
      a[k] = (A + B) + cmul(mk((T)0, (T)1), cmul(conj(tw[k]), A - B));
    }
  }
  __syncthreads();
  const cx<T> *z = stockham_row<+1, T>(a, b, ff, tw, tt, live);
  if (live) {
    cx<T> *out = reinterpret_cast<cx<T> *>(G + rs.gbase[rl]);
    for (int n = tt; n < NC; n += FFT_ANY_TPR) st_nt(out + n, z[n]);
  }
}

// grid -> Fourier (m <= M-1) for any FftFactors length; the 8 columns'
// coefficients leave the block together, m by m, as contiguous runs.
template <class T>
__global__ __launch_bounds__(256) void k_fft_any(const T *__restrict__ G, cx<T> *__restrict__ F, int C,
                                                 ColMap cm, int J, int M, int mode, ActiveMask am,
                                                 FftFactors ff, const cx<T> *__restrict__ tw) {
  extern __shared__ unsigned char fft_any_lds[];
  __shared__ AnyRows rs;
  const int NC = ff.NC, NI = ff.NI;
  const int rl = threadIdx.x / FFT_ANY_TPR, tt = threadIdx.x - rl * FFT_ANY_TPR;
  cx<T> *a = reinterpret_cast<cx<T> *>(fft_any_lds) + (size_t)rl * 2 * NC, *b = a + NC;
  cx<T> *twl = reinterpret_cast<cx<T> *>(fft_any_lds) + (size_t)FFT_ANY_ROWS * 2 * NC;
  for (int k = threadIdx.x; k < NI; k += blockDim.x) twl[k] = tw[k];
  any_rows(rs, C, cm, J, NI, am);
  __syncthreads();
  tw = twl;
  const bool live = rs.fbase[rl] >= 0;
  if (live) {
    const cx<T> *in = reinterpret_cast<const cx<T> *>(G + rs.gbase[rl]);
    for (int n = tt; n < NC; n += FFT_ANY_TPR) a[n] = ld_nt(in + n);
  }
  __syncthreads();
  const cx<T> *Z = stockham_row<-1, T>(a, b, ff, tw, tt, live);
  // X_k = (s - i w d) / 2 / NI, s = Z_k + conj Z_{NC-k}, d = Z_k - conj Z_{NC-k},
  // w = e^{-2 pi i k / NI} (spectral.py:159-164), into the other buffer
  cx<T> *X = Z == a ? b : a;
  if (live) {
    const T half = (T)0.5;
    for (int k = tt; k < M; k += FFT_ANY_TPR) {
      const cx<T> zk = Z[k], zc = conj(Z[(NC - k) % NC]);
      const cx<T> s = zk + zc, d = zk - zc;
      const cx<T> wd = cmul(tw[k], d);
      const T sc = mode == 2 ? fourier_scale_any<T>(2, k, NI) : (T)1 / (T)NI;
      X[k] = sc * mk(half * (s.x + wd.y), half * (s.y - wd.x));
    }
  }
  __syncthreads();
  const size_t ms = (size_t)J * fpad(C);
  const int xoff = (int)(X - a);  // X is the same buffer (0 or NC) of every row
  for (int idx = threadIdx.x; idx < FFT_ANY_ROWS * M; idx += blockDim.x) {
    const int q = idx & (FFT_ANY_ROWS - 1), m = idx / FFT_ANY_ROWS;
    if (rs.fbase[q] >= 0)
      F[rs.fbase[q] + (size_t)m * ms] = reinterpret_cast<cx<T> *>(fft_any_lds)[(size_t)q * 2 * NC + xoff + m];
  }
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
struct SynHist {  // (vor, div, t, z) -> (psi, chi, u, v, t, z): the history fields of model.py:159-165
  static constexpr int NIN = 4, NOUT = 6;
  static constexpr bool hasA(int) { return true; }
  static constexpr bool hasB(int o) { return o == 2 || o == 3; }
  static constexpr bool scaled(int o) { return o == 2 || o == 3; }
  template <class T>
  __device__ __forceinline__ static void terms(const cx<T> *x, T il, T fm, cx<T> *A, cx<T> *B) {
    const cx<T> psi = il * x[0], chi = il * x[1];
    A[0] = psi;
    A[1] = chi;
    A[2] = im_mul(fm, chi);  // u = P (im chi) - H psi  (vordiv_to_uv, spectral.py:255-276)
    B[2] = -psi;
    A[3] = im_mul(fm, psi);  // v = P (im psi) + H chi
    B[3] = chi;
    A[4] = x[2];
    A[5] = x[3];
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
// The tendency analysis (FU, FV, KE, XU, XV, GT, YU, YV, G) -> (vor, div, t,
// q) of dynamics.py:219,246-262 in three groups that share no input: three
// launches of three fields each need fewer registers per wave (more waves
// and loads in flight) than one of nine.
struct AnaTendVD {  // (FU, FV, KE) -> (vor, div): dynamics.py:246-250
  static constexpr int NIN = 3, NOUT = 2;
  static constexpr bool ACC = false;
  static constexpr bool needP(int) { return true; }
  static constexpr bool needH(int f) { return f < 2; }
  template <class T>
  __device__ __forceinline__ static void combine(const cx<T> *AP, const cx<T> *AH, T fm, T nn1a2, cx<T> *o) {
    o[0] = im_mul(fm, AP[1]) + AH[0];                  // curl(fu, fv)
    o[1] = im_mul(fm, AP[0]) - AH[1] + nn1a2 * AP[2];  // div(fu, fv) - lap(ke)
  }
};
struct AnaFluxDiv {  // (XU, XV, GT) -> t or (YU, YV, G) -> q: dynamics.py:219,258-262
  static constexpr int NIN = 3, NOUT = 1;
  static constexpr bool ACC = false;
  static constexpr bool needP(int f) { return f != 1; }
  static constexpr bool needH(int f) { return f == 1; }
  template <class T>
  __device__ __forceinline__ static void combine(const cx<T> *AP, const cx<T> *AH, T fm, T, cx<T> *o) {
    o[0] = AP[2] - im_mul(fm, AP[0]) + AH[1];          // -div(xu, xv) + gt
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
