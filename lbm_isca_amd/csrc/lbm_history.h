// Time-varying forcing, history extraction and spectral resizing.
#pragma once

#include "lbm_model.h"

namespace lbm {

// a + ((b - a) / nt) k as numpy evaluates forcing.py:151-153: the slope
// ka = (b - a) / nt in the fields' precision U, then a + ka k in float64
// (numpy promotes ka times the int64 arange to float64), each operation
// rounded on its own (no fused multiply-add)
template <class U>
__device__ inline double mjo_lerp(U a, U b, U nt, int k) {
#pragma clang fp contract(off)
  const U ka = (b - a) / nt;
  const double x = (double)ka * (double)k;
  return (double)a + x;
}

// Current MJO temperature forcing (forcing.py:99-161), from the step
// counter so that it can sit inside a captured graph.  Call c (1-based,
// one per dynamics call, start-up steps included) is call0 + step + 1.
// The first spinup - 1 calls hold day 1; from call spinup on, day
// it = 1 + q / nt with q = c - spinup is interpolated toward day it + 1 over
// nt calls, k = q mod nt, both days clamped to ndays.  single: the daily
// fields are float32 values (a file's real*4) and the slope is rounded to
// float32, as the reference does with them.
// days: [f][day][j][row][I]; out: [f][j][row][I]; n = J * rows * I.
template <class T>
__global__ void k_mjo(const T *__restrict__ days, T *__restrict__ out, const Ctl *ctl, int nf,
                      int ndays, int nt, int spinup, int call0, int single, size_t n) {
  const long c = (long)call0 + ctl->step + 1;
  int ia = 0, ib = 0, k = 0;
  if (c >= spinup) {
    const long q = c - spinup;
    const long it = 1 + q / nt;
    k = (int)(q % nt);
    ia = (int)(std::min<long>(it, ndays) - 1);
    ib = (int)(std::min<long>(it + 1, ndays) - 1);
  }
  const size_t tot = (size_t)nf * n;
  for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < tot;
       g += (size_t)gridDim.x * blockDim.x) {
    const size_t f = g / n, r = g - f * n;
    const T *d = days + f * (size_t)ndays * n + r;
    const T a = d[(size_t)ia * n];
    const T b = d[(size_t)ib * n];
    out[g] = (T)(single ? mjo_lerp<float>((float)a, (float)b, (float)nt, k)
                        : mjo_lerp<double>((double)a, (double)b, (double)nt, k));
  }
}

// History spectral slot [s][3][L][E] (vor, div, t) times scale -> the
// synthesis input [s][4][L][E] (vor, div, t, z) with the geopotential height
// z = (WW t) / g, (WW t)_l = sum_{k<l} R dlog3_k t_k + R alpha_l t_l
// (vertical.py:196-203, model.py:163).  clear: zero the slot afterwards.
template <class T>
__global__ void k_hist_spec_out(cx<T> *__restrict__ slot, cx<T> *__restrict__ out,
                                const T *__restrict__ vc, int E, int L, int nspec, T scale,
                                int clear) {
  const size_t tot = (size_t)nspec * E;
  const VOff vo{L};
  const T rg = (T)R_DRY, rgrav = (T)1 / (T)GRAV;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot;
       q += (size_t)gridDim.x * blockDim.x) {
    const size_t s = q / E, e = q - s * E;
    const size_t EL = (size_t)E * L;
    cx<T> acc = zero<T>();
    for (int l = 0; l < L; ++l) {
      const size_t iv = s * 3 * EL + (size_t)l * E + e;
      const cx<T> vor = scale * slot[iv], div = scale * slot[iv + EL], t = scale * slot[iv + 2 * EL];
      const size_t o = s * 4 * EL + (size_t)l * E + e;
      out[o] = vor;
      out[o + EL] = div;
      out[o + 2 * EL] = t;
      out[o + 3 * EL] = rgrav * (acc + (rg * vc[vo.alpha() + l]) * t);
      acc = acc + (rg * vc[vo.dlog3() + l]) * t;
      if (clear) slot[iv] = slot[iv + EL] = slot[iv + 2 * EL] = zero<T>();
    }
  }
}

// The last top sample [s][2][L][E] (vor, div) and the current spectral state
// [s][4][L][E] -> synthesis input (vor_last, div_last, t, z) of
// LBM.diagnostics (model.py:150-166: psi, chi from the last tendencies, the
// rest from the current state).
template <class T>
__global__ void k_diag_spec(const cx<T> *__restrict__ last, const cx<T> *__restrict__ sa3,
                            cx<T> *__restrict__ out, const T *__restrict__ vc, int E, int L,
                            int nspec) {
  const size_t tot = (size_t)nspec * E;
  const VOff vo{L};
  const T rg = (T)R_DRY, rgrav = (T)1 / (T)GRAV;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot;
       q += (size_t)gridDim.x * blockDim.x) {
    const size_t s = q / E, e = q - s * E;
    const size_t EL = (size_t)E * L;
    cx<T> acc = zero<T>();
    for (int l = 0; l < L; ++l) {
      const size_t o = s * 4 * EL + (size_t)l * E + e;
      const size_t il = s * 2 * EL + (size_t)l * E + e;
      const cx<T> t = sa3[o + 2 * EL];
      out[o] = last[il];
      out[o + EL] = last[il + EL];
      out[o + 2 * EL] = t;
      out[o + 3 * EL] = rgrav * (acc + (rg * vc[vo.alpha() + l]) * t);
      acc = acc + (rg * vc[vo.dlog3() + l]) * t;
    }
  }
}

// Grid history slot [e][j][L+1][I] times scale -> w (E, J, I, L) and
// p (E, J, I) float64 (the reference layout); clear: zero the slot.
template <class T>
__global__ void k_hist_grid_out(T *__restrict__ slot, double *__restrict__ w, double *__restrict__ p,
                                int E, int J, int I, int L, T scale, int clear) {
  const int R = L + 1;
  const size_t tot = (size_t)E * J * R * I;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot;
       q += (size_t)gridDim.x * blockDim.x) {
    // q runs over the output order (e, j, i, row) so that the level-fastest
    // stores of w are contiguous
    const int r = (int)(q % R);
    size_t rest = q / R;
    const int i = (int)(rest % I);
    rest /= I;
    const int j = (int)(rest % J);
    const int e = (int)(rest / J);
    const size_t src = (((size_t)e * J + j) * R + r) * I + i;
    const double v = (double)(scale * slot[src]);
    if (r < L) {
      if (w) w[(((size_t)e * J + j) * I + i) * L + r] = v;
    } else if (p) {
      p[((size_t)e * J + j) * I + i] = v;
    }
    if (clear) slot[src] = (T)0;
  }
}

// Steady vorticity, divergence and ln ps forcings (model.py:202-205,
// dynamics.py:189-190,246-250): their analyses, constant in time, added to
// the analysed tendencies of vorticity, divergence and pi.  sfv [s][2][L][nf]
// (or null), sfp [s][nf] (or null); one thread per (s, member).
template <class T>
__global__ void k_add_steady(cx<T> *__restrict__ tend3, cx<T> *__restrict__ tend2,
                             const cx<T> *__restrict__ sfv, const cx<T> *__restrict__ sfp, int E,
                             int L, int nspec, int nf, int ecnt) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t s = idx / ecnt, e = idx - s * ecnt;
  if (s >= (size_t)nspec) return;
  const size_t ef = nf > 1 ? e : 0, EL = (size_t)E * L, FL = (size_t)nf * L;
  if (sfv)
    for (int f = 0; f < 2; ++f)
      for (int l = 0; l < L; ++l) {
        cx<T> &x = tend3[(s * 4 + f) * EL + (size_t)l * E + e];
        x = x + sfv[(s * 2 + f) * FL + (size_t)l * nf + ef];
      }
  if (sfp) {
    cx<T> &x = tend2[s * 2 * E + e];
    x = x + sfp[s * nf + ef];
  }
}

// Spectral (B, Ms, Ms, K) -> (B, Md, Md, K) complex128: zero-padding or
// dropping the high-n tail (regrid.py:68-77).
__global__ void k_spec_resize(const double2 *__restrict__ src, double2 *__restrict__ dst, int B,
                              int Ms, int Md, int K) {
  const size_t tot = (size_t)B * Md * Md * K;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot;
       q += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(q % K);
    size_t rest = q / K;
    const int n = (int)(rest % Md);
    rest /= Md;
    const int m = (int)(rest % Md);
    const int b = (int)(rest / Md);
    double2 v = make_double2(0.0, 0.0);
    if (m < Ms && n < Ms) v = src[(((size_t)b * Ms + m) * Ms + n) * K + k];
    dst[q] = v;
  }
}

}  // namespace lbm
