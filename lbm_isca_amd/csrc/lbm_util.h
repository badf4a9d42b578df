// Element-wise operators of the reference's State, Spectral and VerticalGrid
// APIs that act on whole grid or spectral arrays outside the time loop:
// the Robert-Asselin combination (state.py:110-128), the exact pi' <-> ps'
// pair (state.py:77-83), finiteness / max-abs statistics (state.py:131-142),
// the diagonal Laplacian scalings (spectral.py:202-208) and the hydrostatic
// sum phi = Tv WW^T (vertical.py:196-203).  All FP64 on caller buffers.
#pragma once
#include "lbm_model.h"

namespace lbm {

// prev <- cur + gamma (prev - 2 cur + new), n doubles
__global__ void k_robert_asselin(double *__restrict__ prev, const double *__restrict__ cur,
                                 const double *__restrict__ nxt, double gamma, size_t n) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x) {
    const double c = cur[q];
    prev[q] = c + gamma * (prev[q] - 2.0 * c + nxt[q]);
  }
}

// inverse = 0: out = ln(ps_bar + in) - ln(ps_bar) = log1p(in / ps_bar);
// inverse = 1: out = ps_bar (exp(in) - 1) = ps_bar expm1(in).  ps_bar has
// nper values and repeats over the leading axis of in / out.
__global__ void k_lnps(const double *__restrict__ in, const double *__restrict__ psbar,
                       double *__restrict__ out, size_t n, size_t nper, int inverse) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x) {
    const double pb = psbar[q % nper];
    out[q] = inverse ? pb * expm1(in[q]) : log1p(in[q] / pb);
  }
}

// Max of the bit patterns of |x| over finite x (non-negative doubles order
// like their bit patterns) and the number of non-finite x, over this
// block's grid-stride share of x[0, n); reduced in LDS, valid in thread 0.
__device__ __forceinline__ void block_stats(const double *__restrict__ x, size_t n,
                                            unsigned long long &mx_out,
                                            unsigned long long &bad_out) {
  __shared__ unsigned long long smax[256];
  __shared__ unsigned long long sbad[256];
  unsigned long long mx = 0, bad = 0;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x) {
    const double v = x[q];
    if (isfinite(v)) {
      const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(v));
      mx = b > mx ? b : mx;
    } else {
      ++bad;
    }
  }
  smax[threadIdx.x] = mx;
  sbad[threadIdx.x] = bad;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const unsigned long long o = smax[threadIdx.x + w];
      smax[threadIdx.x] = o > smax[threadIdx.x] ? o : smax[threadIdx.x];
      sbad[threadIdx.x] += sbad[threadIdx.x + w];
    }
    __syncthreads();
  }
  mx_out = smax[0];
  bad_out = sbad[0];
}

// stats[0] = bits of max |x| over finite x, stats[1] = number of non-finite
// x; one atomic per block and statistic (stats starts at zero).
__global__ __launch_bounds__(256) void k_grid_stats(const double *__restrict__ x, size_t n,
                                                    unsigned long long *stats) {
  unsigned long long mx, bad;
  block_stats(x, n, mx, bad);
  if (threadIdx.x == 0) {
    atomicMax(&stats[0], mx);
    atomicAdd(&stats[1], bad);
  }
}

// The same statistics without atomics: block b writes its pair to
// part[2b], part[2b+1]; k_grid_stats_final reduces the gridDim.x pairs.
// Nothing has to start at zero.
__global__ __launch_bounds__(256) void k_grid_stats_part(const double *__restrict__ x, size_t n,
                                                         unsigned long long *__restrict__ part) {
  unsigned long long mx, bad;
  block_stats(x, n, mx, bad);
  if (threadIdx.x == 0) {
    part[2 * (size_t)blockIdx.x] = mx;
    part[2 * (size_t)blockIdx.x + 1] = bad;
  }
}

// One block of 256: stats[0..1] from the nb pairs of part.
__global__ __launch_bounds__(256) void k_grid_stats_final(const unsigned long long *__restrict__ part,
                                                          int nb, unsigned long long *__restrict__ stats) {
  __shared__ unsigned long long smax[256];
  __shared__ unsigned long long sbad[256];
  unsigned long long mx = 0, bad = 0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    const unsigned long long m = part[2 * b];
    mx = m > mx ? m : mx;
    bad += part[2 * b + 1];
  }
  smax[threadIdx.x] = mx;
  sbad[threadIdx.x] = bad;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const unsigned long long o = smax[threadIdx.x + w];
      smax[threadIdx.x] = o > smax[threadIdx.x] ? o : smax[threadIdx.x];
      sbad[threadIdx.x] += sbad[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    stats[0] = smax[0];
    stats[1] = sbad[0];
  }
}

// (B, M, M, K) complex: out[b, m, n, k] = fac[n] in[b, m, n, k]
__global__ void k_spec_scale(const double *__restrict__ in, double *__restrict__ out,
                             const double *__restrict__ fac, size_t n, int M, int K) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x) {
    const size_t c = q >> 1;  // complex element
    const int deg = (int)((c / (size_t)K) % (size_t)M);
    out[q] = fac[deg] * in[q];
  }
}

// phi_k = R alpha_k x_k + sum_{l<k} R dlog3_l x_l along the level axis of
// (npts, L, inner) doubles (inner = 1 real, 2 complex interleaved).  One
// thread per (point, component); in place is allowed.
__global__ void k_hydrostatic(const double *in, double *out,
                              const double *__restrict__ alpha, const double *__restrict__ dlog3,
                              size_t npts, int L, int inner) {
  const size_t tot = npts * (size_t)inner;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot;
       q += (size_t)gridDim.x * blockDim.x) {
    const size_t pt = q / inner;
    const int c = (int)(q - pt * inner);
    const double *x = in + pt * L * inner + c;
    double *y = out + pt * L * inner + c;
    double acc = 0.0;
    for (int k = 0; k < L; ++k) {
      const double xk = x[(size_t)k * inner];
      y[(size_t)k * inner] = acc + R_DRY * alpha[k] * xk;
      acc += R_DRY * dlog3[k] * xk;
    }
  }
}

}  // namespace lbm
