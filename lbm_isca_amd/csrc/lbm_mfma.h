// Legendre analysis and synthesis on the CDNA4 matrix cores.
//
// For one zonal wavenumber m both transforms are small dense contractions
// over latitude (analysis) or degree n (synthesis), folded north/south with
// P(-mu) = (-1)^(n+m) P(mu), H(-mu) = -(-1)^(n+m) H(mu).  They are computed
// as 16x16 output tiles with 16x16x4 MFMA steps, real tables times the real
// and imaginary parts of the complex data as separate columns: a tile covers
// 8 complex columns, real column j holding part j>>3 (0 re, 1 im) of column
// j&7.
//
// Operand layout of v_mfma_{f64,f32}_16x16x4 (64 lanes):
//   A[i][k]: lane i + 16k;  B[k][j]: lane j + 16k;
//   D[i][j] value v of lane l: j = l%16, i = l/16 + 4v (f64) or
//   4(l/16) + v (f32).
// Split-BF16x3 (FP32 data, SPLIT kernels): v_mfma_f32_16x16x16bf16_1k with
//   A[i][k]: lane i + 16(k/4), element k%4;  B[k][j]: lane j + 16(k/4),
//   element k%4;  D as f32.  Each operand x = hi + lo (hi = bf16(x), lo =
//   bf16(x - hi)) and a b ~ hi_a hi_b + hi_a lo_b + lo_a hi_b in FP32
//   accumulation (the lo_a lo_b term, ~2^-16 relative, is dropped).  The
//   Legendre tables come pre-split from the plan (SplitTabs); the data
//   operand is split as it is staged.
#pragma once
#include "lbm_device.h"

namespace lbm {

template <class T>
struct Mfma;

template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

struct Bf16x3 {
  typedef float v4 __attribute__((ext_vector_type(4)));
  typedef short s4 __attribute__((ext_vector_type(4)));
  // round to nearest even (finite inputs)
  static __device__ __forceinline__ short rne(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
  }
  static __device__ __forceinline__ float up(short h) {
    return __uint_as_float(((unsigned)(unsigned short)h) << 16);
  }
  static __device__ __forceinline__ void split(const float *x, s4 &hi, s4 &lo) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const short h = rne(x[q]);
      hi[q] = h;
      lo[q] = rne(x[q] - up(h));
    }
  }
  static __device__ __forceinline__ v4 mma3(s4 ah, s4 al, s4 bh, s4 bl, v4 c) {
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, c, 0, 0, 0);
  }
};

// select between two loaded values (keeps aggregates out of private memory)
template <class T>
__device__ __forceinline__ T pick(T re, T im, int part) {
  return part == 0 ? re : im;
}

// complex value of real column j of a tile from the lane's own part and the
// partner lane j^8 (same k-group)
template <class T>
__device__ inline cx<T> pair_of(T mine, int part) {
  const T other = __shfl_xor(mine, 8);
  return part == 0 ? mk(mine, other) : mk(other, mine);
}

// ---------------------------------------------------------------------------
// Analysis.  Block = 4 waves for (parity p = w&1, n-tile t = w>>1) of one m
// and 8 columns; n = m + p + 2(16t + i); blockIdx.z selects n-tiles 2z,
// 2z+1 (more than one only when N+1-m > 64).  K = latitude pairs, staged 32
// at a time per field, double-buffered in LDS.
// ---------------------------------------------------------------------------
// whether any input field of an analysis recipe contracts with P (H)
template <class R>
constexpr bool any_p() {
  for (int f = 0; f < R::NIN; ++f)
    if (R::needP(f)) return true;
  return false;
}
template <class R>
constexpr bool any_h() {
  for (int f = 0; f < R::NIN; ++f)
    if (R::needH(f)) return true;
  return false;
}

// Tables pre-split into BF16 high and low parts (split-BF16x3 plans only):
// the same [s][Jh] layout as the FP32 table they come from.
struct SplitTabs {
  const short *Ph = nullptr, *Pl = nullptr, *Hh = nullptr, *Hl = nullptr;
};

template <class T, class R, bool SPLIT = false, int OCC = 4>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC))) void k_ana_mfma(const cx<T> *__restrict__ in,
                                                  cx<T> *__restrict__ out,
                                                  const T *__restrict__ Pa,
                                                  const T *__restrict__ Ha,
                                                  const T *__restrict__ nn1a2, int ncols, int C,
                                                  InSlots slots, int J, int M, int out_nf,
                                                  int out_f0, ActiveMask am, SplitTabs sp) {
  typedef typename Mfma<T>::v4 v4;
  constexpr int NIN = R::NIN, NOUT = R::NOUT;
  // one symmetric/antisymmetric buffer per field: a chunk is stored whole,
  // then contracted (two barriers per chunk)
  __shared__ T sym[NIN][32][16], anti[NIN][32][16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int p = w & 1, tile = (w >> 1) + 2 * blockIdx.z;
  const int m = blockIdx.y, col0 = am.tile_col0(blockIdx.x);
  const int Jh = J / 2;
  const int N = M - 1;
  const int s0 = spec_off(m, M) - m;
  const int nch = (Jh + 31) / 32;
  // the lowest degree of this block's tiles is m + 64 blockIdx.z: a block
  // beyond N has no live wave and exits before any barrier
  if (m + 64 * (int)blockIdx.z > N) return;
  // a tile of finished members only: nothing to do (uniform per block)
  if (!am.any8(col0, ncols)) return;

  // loader mapping: thread -> (latitude pair lj of the chunk, column lc)
  const int lj = t >> 3, lc = t & 7;
  const bool lcol_ok = (col0 + lc) < ncols;
  int off[NIN];
#pragma unroll
  for (int f = 0; f < NIN; ++f) off[f] = slots.off[f];
  // raw north/south values of this thread's (latitude pair, column) for
  // every field: the loads of all fields are in flight together, and the
  // registers of a field already staged take the next chunk's values
  cx<T> rn[NIN], rs[NIN];
  // the Fourier slab of m as a buffer (32-bit per-lane byte offsets)
  const size_t slab = (size_t)J * fpad(C);
  const BufView bin(in + (size_t)m * slab, (unsigned)(slab * sizeof(cx<T>)));
  const cx<T> *const tag = nullptr;
  auto load = [&](int jc, int f) __attribute__((always_inline)) {
    const int jp = 32 * jc + lj;
    rn[f] = rs[f] = zero<T>();
    if (lcol_ok && jp < Jh) {
      const int c = off[f] + col0 + lc;
      rn[f] = bin.ld_nt((unsigned)(fidx(0, jp, c, J, C) * sizeof(cx<T>)), 0u, tag);
      rs[f] = bin.ld_nt((unsigned)(fidx(0, J - 1 - jp, c, J, C) * sizeof(cx<T>)), 0u, tag);
    }
  };
  auto store = [&](int f, int buf) __attribute__((always_inline)) {
    sym[buf][lj][lc] = rn[f].x + rs[f].x;
    sym[buf][lj][8 + lc] = rn[f].y + rs[f].y;
    anti[buf][lj][lc] = rn[f].x - rs[f].x;
    anti[buf][lj][8 + lc] = rn[f].y - rs[f].y;
  };

  // A operand lane (i = l%16 -> n row, k = l/16 -> latitude pair); B lane (j = l%16)
  const int i_a = lane & 15, k_a = lane >> 4;
  const int n_a = m + p + 2 * (16 * tile + i_a);
  const bool row_ok = n_a <= N;
  const bool wave_live = (m + p + 32 * tile) <= N;
  const int j_b = lane & 15;
  const size_t trow = (size_t)(s0 + (row_ok ? n_a : m)) * Jh;

  v4 AP[NIN], AH[NIN];
#pragma unroll
  for (int f = 0; f < NIN; ++f) AP[f] = AH[f] = (v4){0, 0, 0, 0};
  constexpr bool anyP = any_p<R>(), anyH = any_h<R>();

#pragma unroll
  for (int f = 0; f < NIN; ++f) load(0, f);
  for (int jc = 0; jc < nch; ++jc) {
    if (jc > 0) __syncthreads();  // the previous chunk's contractions are done
#pragma unroll
    for (int f = 0; f < NIN; ++f) store(f, f);
    __syncthreads();
    // this chunk is in LDS: the registers load the next one's
    if (jc + 1 < nch) {
#pragma unroll
      for (int f = 0; f < NIN; ++f) load(jc + 1, f);
    }
    // Legendre operands of this chunk, shared by every field: lane row n_a,
    // latitude pair 32 jc + 4 q + k_a (FP64/FP32 K = 4 steps), or, split,
    // the pre-split halves of pairs 32 jc + 16 h + 4 k_a .. +3 (bf16 K = 16
    // steps; Jh is a multiple of 4, so a group is all in range or all out)
    T pr[8], hr[8];
    Bf16x3::s4 prh[2], prl[2], hrh[2], hrl[2];
    if constexpr (SPLIT) {
      const Bf16x3::s4 z4 = {0, 0, 0, 0};
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int jp0 = 32 * jc + 16 * h + 4 * k_a;
        const bool ok = wave_live && row_ok && jp0 < Jh;
        const size_t ti = trow + (ok ? jp0 : 0);
        prh[h] = (anyP && ok) ? *reinterpret_cast<const Bf16x3::s4 *>(sp.Ph + ti) : z4;
        prl[h] = (anyP && ok) ? *reinterpret_cast<const Bf16x3::s4 *>(sp.Pl + ti) : z4;
        hrh[h] = (anyH && ok) ? *reinterpret_cast<const Bf16x3::s4 *>(sp.Hh + ti) : z4;
        hrl[h] = (anyH && ok) ? *reinterpret_cast<const Bf16x3::s4 *>(sp.Hl + ti) : z4;
      }
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int jp = 32 * jc + 4 * q + k_a;
        const bool ok = wave_live && row_ok && jp < Jh;
        pr[q] = (anyP && ok) ? Pa[trow + jp] : (T)0;
        hr[q] = (anyH && ok) ? Ha[trow + jp] : (T)0;
      }
    }
#pragma unroll
    for (int f = 0; f < NIN; ++f) {
      const int buf = f;
      if (wave_live) {
        // the P rows of parity p contract with the symmetric sums when
        // p = 0, the H rows with the other one
        const T(*bP)[16] = p == 0 ? sym[buf] : anti[buf];
        const T(*bH)[16] = p == 0 ? anti[buf] : sym[buf];
        if constexpr (SPLIT) {
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            float xp[4], xh[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int kk = 16 * h + 4 * k_a + q;
              if (R::needP(f)) xp[q] = bP[kk][j_b];
              if (R::needH(f)) xh[q] = bH[kk][j_b];
            }
            Bf16x3::s4 bh, bl;
            if (R::needP(f)) {
              Bf16x3::split(xp, bh, bl);
              AP[f] = Bf16x3::mma3(prh[h], prl[h], bh, bl, AP[f]);
            }
            if (R::needH(f)) {
              Bf16x3::split(xh, bh, bl);
              AH[f] = Bf16x3::mma3(hrh[h], hrl[h], bh, bl, AH[f]);
            }
          }
        } else {
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const int kk = 4 * q + k_a;
            if (R::needP(f)) AP[f] = Mfma<T>::mma(pr[q], bP[kk][j_b], AP[f]);
            if (R::needH(f)) AH[f] = Mfma<T>::mma(hr[q], bH[kk][j_b], AH[f]);
          }
        }
      }
    }
  }
  if (!wave_live) return;
  // combine and store: lane holds rows i = row(lane, v), column j = lane%16
  const int c = j_b & 7, part = j_b >> 3;
  const int col = col0 + c;
  const T fm = (T)m;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int i = Mfma<T>::row(lane, v);
    const int n = m + p + 2 * (16 * tile + i);
    cx<T> ap[NIN], ah[NIN], o[NOUT];
#pragma unroll
    for (int f = 0; f < NIN; ++f) {
      ap[f] = pair_of<T>(AP[f][v], part);
      ah[f] = pair_of<T>(AH[f][v], part);
    }
    if (n > N || col >= ncols || !am(col % am.E)) continue;
    R::template combine<T>(ap, ah, fm, nn1a2[n], o);
    const int s = s0 + n;
#pragma unroll
    for (int q = 0; q < NOUT; ++q) {
      T *dst = reinterpret_cast<T *>(out + ((size_t)s * out_nf + out_f0 + q) * ncols + col);
      if constexpr (R::ACC)
        dst[part] += pick<T>(o[q].x, o[q].y, part);
      else
        dst[part] = pick<T>(o[q].x, o[q].y, part);
    }
  }
}

// ---------------------------------------------------------------------------
// Synthesis.  Block = 4 waves: latitude tile rt = w&1 (pairs 32z + 16rt ..
// +15) and output group g = w>>1 (outputs o with o%2 == g, so each wave
// holds half of the accumulators), for one m and 8 columns.  K = degree n:
// 32 consecutive rows n = m + 32c .. are staged per chunk c, each parity p
// contributing 16 rows n = m + p + 2kk; the recipe terms A_o(n), B_o(n) are
// formed per lane.
// ---------------------------------------------------------------------------
template <class T, class R, int G, bool SPLIT>
__device__ __forceinline__ void syn_group(const cx<T> (*xin)[32][8], bool first_chunk, int r0, int nn, int m,
                                 int s0, int Jh, int jp_a, bool jp_ok, int lane,
                                 const T *__restrict__ Ps, const T *__restrict__ Hs,
                                 const SplitTabs &sp, const T *__restrict__ inv_lap,
                                 typename Mfma<T>::v4 (&EA)[R::NOUT], typename Mfma<T>::v4 (&OA)[R::NOUT],
                                 typename Mfma<T>::v4 (&EB)[R::NOUT], typename Mfma<T>::v4 (&OB)[R::NOUT]) {
  constexpr int NIN = R::NIN, NOUT = R::NOUT;
  const int k_a = lane >> 4, j_b = lane & 15;
  const int c_b = j_b & 7, part_b = j_b >> 3;
  const T fm = (T)m;
  if constexpr (SPLIT) {
    // one K = 16 step per parity: rows p + 2 (4 (lane/16) + q), q < 4
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      Bf16x3::s4 ph, pl, hh, hl, bh, bl;
      float bA[NOUT][4], bB[NOUT][4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int ra = p + 2 * (4 * k_a + q);
        const bool a_ok = jp_ok && (r0 + ra) < nn;
        const size_t ti = (size_t)(s0 + (a_ok ? m + r0 + ra : m)) * Jh + (jp_ok ? jp_a : 0);
        ph[q] = a_ok ? sp.Ph[ti] : (short)0;
        pl[q] = a_ok ? sp.Pl[ti] : (short)0;
        hh[q] = a_ok ? sp.Hh[ti] : (short)0;
        hl[q] = a_ok ? sp.Hl[ti] : (short)0;
        cx<T> x[NIN], A[NOUT], B[NOUT];
#pragma unroll
        for (int o = 0; o < NOUT; ++o) A[o] = B[o] = zero<T>();
        if (r0 + ra < nn) {
#pragma unroll
          for (int f = 0; f < NIN; ++f) x[f] = xin[f][ra][c_b];
          R::template terms<T>(x, inv_lap[m + r0 + ra], fm, A, B);
        }
#pragma unroll
        for (int o = G; o < NOUT; o += 2) {
          bA[o][q] = (float)pick<T>(A[o].x, A[o].y, part_b);
          bB[o][q] = (float)pick<T>(B[o].x, B[o].y, part_b);
        }
      }
#pragma unroll
      for (int o = G; o < NOUT; o += 2) {
        if (R::hasA(o)) {
          Bf16x3::split(bA[o], bh, bl);
          if (p == 0)
            EA[o] = Bf16x3::mma3(ph, pl, bh, bl, EA[o]);
          else
            OA[o] = Bf16x3::mma3(ph, pl, bh, bl, OA[o]);
        }
        if (R::hasB(o)) {
          Bf16x3::split(bB[o], bh, bl);
          if (p == 0)
            EB[o] = Bf16x3::mma3(hh, hl, bh, bl, EB[o]);
          else
            OB[o] = Bf16x3::mma3(hh, hl, bh, bl, OB[o]);
        }
      }
    }
    (void)first_chunk;
    return;
  }
#pragma unroll
  for (int p = 0; p < 2; ++p) {
#pragma unroll
    for (int ks = 0; ks < 16; ks += 4) {
      const int ra = p + 2 * (ks + k_a);
      const bool a_ok = jp_ok && (r0 + ra) < nn;
      const size_t ti = (size_t)(s0 + (a_ok ? m + r0 + ra : m)) * Jh + (jp_ok ? jp_a : 0);
      const T pa = a_ok ? Ps[ti] : (T)0;
      const T ha = a_ok ? Hs[ti] : (T)0;
      const int rb = p + 2 * (ks + (lane >> 4));
      cx<T> x[NIN], A[NOUT], B[NOUT];
#pragma unroll
      for (int o = 0; o < NOUT; ++o) A[o] = B[o] = zero<T>();
      if (r0 + rb < nn) {
#pragma unroll
        for (int f = 0; f < NIN; ++f) x[f] = xin[f][rb][c_b];
        R::template terms<T>(x, inv_lap[m + r0 + rb], fm, A, B);
      }
#pragma unroll
      for (int o = G; o < NOUT; o += 2) {
        if (R::hasA(o)) {
          const T b = pick<T>(A[o].x, A[o].y, part_b);
          if (p == 0)
            EA[o] = Mfma<T>::mma(pa, b, EA[o]);
          else
            OA[o] = Mfma<T>::mma(pa, b, OA[o]);
        }
        if (R::hasB(o)) {
          const T b = pick<T>(B[o].x, B[o].y, part_b);
          if (p == 0)
            EB[o] = Mfma<T>::mma(ha, b, EB[o]);
          else
            OB[o] = Mfma<T>::mma(ha, b, OB[o]);
        }
      }
    }
  }
  (void)first_chunk;
}

template <class T, class R, int G, bool SPLIT>
__device__ __forceinline__ void syn_body(const cx<T> *__restrict__ in, cx<T> *__restrict__ out,
                                const T *__restrict__ Ps, const T *__restrict__ Hs,
                                const SplitTabs &sp, const T *__restrict__ inv_lap, const T *__restrict__ rcos, int ncols,
                                int C, const OutSlots &slots, int J, int M, int in_nf,
                                cx<T> (*xin)[32][8], int col0) {
  typedef typename Mfma<T>::v4 v4;
  constexpr int NIN = R::NIN, NOUT = R::NOUT;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int rt = w & 1;
  const int m = blockIdx.y;
  const int Jh = J / 2;
  const int nn = M - m;
  const int s0 = spec_off(m, M) - m;
  const int jp_a = 32 * blockIdx.z + 16 * rt + (lane & 15);
  const bool jp_ok = jp_a < Jh;
  v4 EA[NOUT], OA[NOUT], EB[NOUT], OB[NOUT];
#pragma unroll
  for (int o = 0; o < NOUT; ++o) EA[o] = OA[o] = EB[o] = OB[o] = (v4){0, 0, 0, 0};
  for (int r0 = 0; r0 < nn; r0 += 32) {
    __syncthreads();
    for (int q = t; q < 32 * 8; q += 256) {
      const int r = q >> 3, c = q & 7;
      const bool ok = (col0 + c) < ncols && (r0 + r) < nn;
#pragma unroll
      for (int f = 0; f < NIN; ++f)
        xin[f][r][c] = ok ? in[((size_t)(s0 + m + r0 + r) * in_nf + f) * ncols + col0 + c] : zero<T>();
    }
    __syncthreads();
    syn_group<T, R, G, SPLIT>(xin, r0 == 0, r0, nn, m, s0, Jh, jp_a, jp_ok, lane, Ps, Hs, sp,
                              inv_lap, EA, OA, EB, OB);
  }
  const int j_b = lane & 15;
  const int c_b = j_b & 7, part_b = j_b >> 3;
  const int col = col0 + c_b;
  if (col >= ncols) return;  // the partner lane (lane ^ 8) has the same column
  // Lanes j and j ^ 8 hold the real and imaginary parts of one column: they
  // swap halves so that each stores whole complex values, the real-part
  // lanes rows v0, the imaginary-part lanes rows v0 + 1.
#pragma unroll
  for (int v0 = 0; v0 < 4; v0 += 2) {
    const int jp0 = 32 * blockIdx.z + 16 * rt + Mfma<T>::row(lane, v0);
    const int jp1 = 32 * blockIdx.z + 16 * rt + Mfma<T>::row(lane, v0 + 1);
    const T sc0 = jp0 < Jh ? rcos[jp0] : (T)0, sc1 = jp1 < Jh ? rcos[jp1] : (T)0;
    const int jp = part_b == 0 ? jp0 : jp1;  // the row this lane stores
#pragma unroll
    for (int o = G; o < NOUT; o += 2) {
      T n0 = EA[o][v0] + OA[o][v0] + EB[o][v0] + OB[o][v0];
      T s0 = EA[o][v0] - OA[o][v0] + OB[o][v0] - EB[o][v0];
      T n1 = EA[o][v0 + 1] + OA[o][v0 + 1] + EB[o][v0 + 1] + OB[o][v0 + 1];
      T s1 = EA[o][v0 + 1] - OA[o][v0 + 1] + OB[o][v0 + 1] - EB[o][v0 + 1];
      if (R::scaled(o)) {
        n0 *= sc0;
        s0 *= sc0;
        n1 *= sc1;
        s1 *= sc1;
      }
      const T gn = __shfl_xor(part_b == 0 ? n1 : n0, 8);
      const T gs = __shfl_xor(part_b == 0 ? s1 : s0, 8);
      if (jp >= Jh) continue;
      const cx<T> zn = part_b == 0 ? mk(n0, gn) : mk(gn, n1);
      const cx<T> zs = part_b == 0 ? mk(s0, gs) : mk(gs, s1);
      cx<T> *pn = out + fidx(m, jp, slots.off[o] + col, J, C);
      cx<T> *ps = out + fidx(m, J - 1 - jp, slots.off[o] + col, J, C);
      if constexpr (sizeof(T) == 8) {
        // streaming past the caches (measured: FP32 loses)
        st_nt(pn, zn);
        st_nt(ps, zs);
      } else {
        *pn = zn;
        *ps = zs;
      }
    }
  }
}

template <class T, class R, bool SPLIT = false, int OCC = 3>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC))) void k_syn_mfma(const cx<T> *__restrict__ in,
                                                  cx<T> *__restrict__ out,
                                                  const T *__restrict__ Ps,
                                                  const T *__restrict__ Hs,
                                                  const T *__restrict__ inv_lap,
                                                  const T *__restrict__ rcos, int ncols, int C,
                                                  OutSlots slots, int J, int M, int in_nf,
                                                  ActiveMask am, SplitTabs sp) {
  __shared__ cx<T> xin[R::NIN][32][8];
  const int col0 = am.tile_col0(blockIdx.x);
  if (!am.any8(col0, ncols)) return;
  if ((threadIdx.x >> 7) == 0)
    syn_body<T, R, 0, SPLIT>(in, out, Ps, Hs, sp, inv_lap, rcos, ncols, C, slots, J, M, in_nf, xin,
                             col0);
  else
    syn_body<T, R, 1, SPLIT>(in, out, Ps, Hs, sp, inv_lap, rcos, ncols, C, slots, J, M, in_nf, xin,
                             col0);
}

}  // namespace lbm
