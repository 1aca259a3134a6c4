// Device kernels of a micrograph's preparation, shared by the search (tm_core.hip) and the
// refinement (tm_refine.hip): ReplaceOutliersWithMean(5), the scaled forward transform with
// ZeroCentralPixel, the whitening curve (Compute1DPowerSpectrumCurve, SquareRoot, Reciprocal,
// normalised to a maximum of 1), ApplyCurveFilter and the normalisation by the spectrum's sum of
// squares, with the reference's float expressions; sums run in FP64 (the reference sums the curve
// bins in float, in scan order). Kernels have internal linkage (one copy per translation unit).
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#include "tm_kernels.hpp"

namespace tmp {
namespace {

using tmk::cf;
using tmk::cmake;


// FP64 sum and sum of squares (of powf(v, 2) in float) of a dense image, per block of a fixed chunk.
template <int NT>
__global__ __launch_bounds__(NT) void k_moments(const float* __restrict__ img, long long n, long long chunk,
                                                double* __restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double r1[NT], r2[NT];
  const long long lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
  double a = 0.0, b = 0.0;
  for (long long e = lo + threadIdx.x; e < hi; e += NT) {
    const float v = img[e];
    a += v;
    b += v * v;
  }
  r1[threadIdx.x] = a;
  r2[threadIdx.x] = b;
  __syncthreads();
  for (int k = NT / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      r1[threadIdx.x] += r1[threadIdx.x + k];
      r2[threadIdx.x] += r2[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r1[0];
    part[2 * blockIdx.x + 1] = r2[0];
  }
}

// mean and the 5-sigma bounds of ReplaceOutliersWithMean (the host expression order)
__global__ void k_outlier_bounds(const double* __restrict__ part, int nparts, long long n, float* __restrict__ st) {
  double s = 0.0, s2 = 0.0;
  for (int q = 0; q < nparts; ++q) {
    s += part[2 * q];
    s2 += part[2 * q + 1];
  }
  const float m = (float)(s / n);
  const float sigma = sqrtf(fabsf(float(s2 / n - (double)(m * m))));
  st[0] = m;
  st[1] = m - 5.0f * sigma;
  st[2] = m + 5.0f * sigma;
}

// dense micrograph -> padded rows (mx + 2) with outliers replaced by the mean
__global__ void k_outlier_pad(const float* __restrict__ src, int mx, int my, const float* __restrict__ st,
                              float* __restrict__ dst) {
  const long long n = (long long)(mx + 2) * my;
  const float m = st[0], lo = st[1], hi = st[2];
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < n; e += (long long)gridDim.x * 256) {
    const int x = (int)(e % (mx + 2));
    const long long y = e / (mx + 2);
    float v = 0.f;
    if (x < mx) {
      v = src[y * mx + x];
      if (v > hi || v < lo) v = m;
    }
    dst[e] = v;
  }
}

__global__ void k_scale_zero_dc(cf* __restrict__ c, long long n, float s) {
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < n; e += (long long)gridDim.x * 256) {
    cf v = c[e];
    v.x *= s;
    v.y *= s;
    if (e == 0) v = cmake(0.f, 0.f);
    c[e] = v;
  }
}

// Curve::AddValueAtXUsingLinearInterpolation weights of one value at wx (curve x[0..n-1])
__device__ inline void curve_add(const float* __restrict__ cx, int n, float wx, float val, double* __restrict__ acc) {
#pragma clang fp contract(off)
  // the reference asserts 0 <= index < n (curve.cpp:528); the clamp keeps a rounding edge in range
  const int idx = min(max((cx[0] == cx[1]) ? n - 2 : int((wx - cx[0]) / (cx[1] - cx[0])), 0), n - 1);
  if (idx == n - 1) {
    atomicAdd(acc + idx, (double)val);
    return;
  }
  const float d = (wx - cx[idx]) / (cx[idx + 1] - cx[idx]);
  const float vm = val * (1.0 - d), vn = val * d;
  atomicAdd(acc + idx, (double)vm);
  atomicAdd(acc + idx + 1, (double)vn);
}

// Compute1DPowerSpectrumCurve (image.cpp:7135-7205) on a W x H half spectrum: intensities and
// counts into the curve's bins (Hermitian mates of the i = 0 column skipped).
__global__ void k_power_curve(const cf* __restrict__ c, int W, int H, const float* __restrict__ cx, int n,
                              double* __restrict__ acc_w, double* __restrict__ acc_n) {
#pragma clang fp contract(off)
  const int nh = W / 2 + 1, fny = H / 2;
  const long long total = (long long)nh * H;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
    const int i = (int)(e % nh), j = (int)(e / nh);
    if (i == 0 && j >= fny) continue;
    const float fy = (float)((j >= fny ? j - H : j) * (1.0 / double(H)));
    const float fx = (float)(i * (1.0 / double(W)));
    const float sf = sqrtf(fx * fx + fy * fy + 0.0f);
    const cf v = c[e];
    curve_add(cx, n, sf, v.x * v.x + v.y * v.y, acc_w);
    curve_add(cx, n, sf, 1.0f, acc_n);
  }
}

// k_power_curve with each block's bins privatised in LDS (dynamic shared memory of 2 n doubles):
// the block accumulates with LDS atomics, then adds each non-zero bin to the global curve once.
__global__ __launch_bounds__(256) void k_power_curve_lds(const cf* __restrict__ c, int W, int H,
                                                         const float* __restrict__ cx, int n,
                                                         double* __restrict__ acc_w, double* __restrict__ acc_n) {
#pragma clang fp contract(off)
  extern __shared__ double bins[];
  for (int q = threadIdx.x; q < 2 * n; q += 256) bins[q] = 0.0;
  __syncthreads();
  const int nh = W / 2 + 1, fny = H / 2;
  const long long total = (long long)nh * H;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
    const int i = (int)(e % nh), j = (int)(e / nh);
    if (i == 0 && j >= fny) continue;
    const float fy = (float)((j >= fny ? j - H : j) * (1.0 / double(H)));
    const float fx = (float)(i * (1.0 / double(W)));
    const float sf = sqrtf(fx * fx + fy * fy + 0.0f);
    const cf v = c[e];
    curve_add(cx, n, sf, v.x * v.x + v.y * v.y, bins);
    curve_add(cx, n, sf, 1.0f, bins + n);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < n; q += 256) {
    if (bins[q] != 0.0) atomicAdd(acc_w + q, bins[q]);
    if (bins[n + q] != 0.0) atomicAdd(acc_n + q, bins[n + q]);
  }
}

// The power curve into acc_w / acc_n (zeroed by the caller): LDS-privatised bins up to 4000 bins
// (64 KB), else one global atomic per contribution.
inline void power_curve(hipStream_t st, const cf* spec, int W, int H, const float* cx, int n, double* acc_w,
                        double* acc_n, int grid) {
  // the LDS variant zeroes and flushes 2 n bins per block: at most 1024 blocks (4 per CU), so that
  // each block's grid-stride loop covers many spectrum elements
  if (n <= 4000)
    hipLaunchKernelGGL(k_power_curve_lds, dim3(std::min(grid, 1024)), dim3(256), 2 * (size_t)n * sizeof(double), st,
                       spec, W, H, cx, n, acc_w, acc_n);
  else
    hipLaunchKernelGGL(k_power_curve, dim3(grid), dim3(256), 0, st, spec, W, H, cx, n, acc_w, acc_n);
}

// whitening filter: mean power per bin, SquareRoot, Reciprocal, scaled to a maximum of 1. One block
// of 256 threads: each bin's value is independent, and the maximum is exact in any order (fmaxf
// skips NaN as the sequential loop does), so the result equals a one-thread loop bit for bit.
__global__ __launch_bounds__(256) void k_curve_finish(const double* __restrict__ acc_w, const double* __restrict__ acc_n,
                                                      int n, float* __restrict__ cy) {
  __shared__ float red[256];
  float mx = -FLT_MAX;
  for (int q = threadIdx.x; q < n; q += 256) {
    const float w = (float)acc_w[q], c = (float)acc_n[q];
    float a = c > 0.0f ? w / c : 0.0f;
    a = sqrtf(a);
    if (a != 0.0f) a = (float)(1.0 / a);
    cy[q] = a;
    mx = fmaxf(mx, a);
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  const float k = 1.0f / red[0];
  for (int q = threadIdx.x; q < n; q += 256) cy[q] *= k;
}

// Curve::ReturnLinearInterpolationFromX
__device__ inline float curve_interp(const float* __restrict__ cx, const float* __restrict__ cy, int n, float wx) {
#pragma clang fp contract(off)
  int i = n - 1;
  if (wx < cx[0]) {
    i = 0;
  } else {
    int c = int((wx - cx[0]) / (cx[1] - cx[0]));
    c = max(0, min(c, n - 1));
    while (c > 0 && wx < cx[c]) --c;
    while (c < n - 1 && wx >= cx[c + 1]) ++c;
    i = c;
  }
  if (i == n - 1) return cy[n - 1];
  float r = 0.0f;
  const float d = (wx - cx[i]) / (cx[i + 1] - cx[i]);
  r += (float)(cy[i] * (1.0 - d));
  r += cy[i + 1] * d;
  return r;
}

// ApplyCurveFilter (image.cpp:7284-7313, resolution limit 1), then ZeroCentralPixel
__global__ void k_apply_curve(cf* __restrict__ c, int W, int H, const float* __restrict__ cx,
                              const float* __restrict__ cy, int n) {
#pragma clang fp contract(off)
  const int nh = W / 2 + 1;
  const int fny = (H % 2 == 0) ? H / 2 : H / 2 + 1;
  const long long total = (long long)nh * H;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
    const int i = (int)(e % nh), j = (int)(e / nh);
    // std::pow of the double frequency, then rounded to float (image.cpp:7298-7303)
    const double fy = (j >= fny ? j - H : j) * (1.0 / double(H));
    const double fx = i * (1.0 / double(W));
    const float s2 = (float)(fx * fx) + (float)(fy * fy) + 0.0f;
    const float sf = (float)sqrt((double)s2);
    cf v = c[e];
    if (e == 0) {
      v = cmake(0.f, 0.f);
    } else if (sf <= 1.0f) {
      const float f = curve_interp(cx, cy, n, sf);
      v.x *= f;
      v.y *= f;
    } else {
      v = cmake(0.f, 0.f);
    }
    c[e] = v;
  }
}

// ReturnSumOfSquares, Fourier branch (image.cpp:290-311) of a 2-D half spectrum: FP64 partials
template <int NT>
__global__ __launch_bounds__(NT) void k_spectrum_ssq(const cf* __restrict__ c, int W, int H, long long chunk,
                                                     double* __restrict__ part) {
  __shared__ double r[NT];
  const int nh = W / 2 + 1;
  const bool xe = W % 2 == 0, ye = H % 2 == 0;
  const int fny = ye ? H / 2 : H / 2 + 1;
  const int lo_y = ye ? -H / 2 : -(H - 1) / 2;
  const long long total = (long long)nh * H;
  const long long lo = blockIdx.x * chunk, hi = min(total, lo + chunk);
  double a = 0.0;
  for (long long e = lo + threadIdx.x; e < hi; e += NT) {
    const int i = (int)(e % nh), j = (int)(e / nh);
    const int jj = j >= fny ? j - H : j;
    const cf v = c[e];
    const double m = (double)hypotf(v.x, v.y);  // |c| in float, squared in double
    const double a2 = m * m;
    const bool edge_x = (i == 0 || (i == W / 2 && xe));
    if (edge_x && (jj == 0 || (jj == lo_y && xe))) a += a2 * 0.5;
    else if ((i != 0 && (i != W / 2 || !xe)) || jj >= 0) a += a2;
  }
  r[threadIdx.x] = a;
  __syncthreads();
  for (int k = NT / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) r[threadIdx.x] += r[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = r[0];
}

// c *= 1 / sqrt(sum of squares): DivideByConstant(sqrt(ReturnSumOfSquares())) of
// refine_template.cpp:544 (float_sqrt 0: the root in double) or the search's input normalisation
// (float_sqrt 1: 1 / sqrtf, match_template.cpp:718-724)
__global__ void k_normalise_spectrum(cf* __restrict__ c, long long n, const double* __restrict__ part, int nparts,
                                     int float_sqrt) {
  double s = 0.0;
  for (int q = 0; q < nparts; ++q) s += part[q];
  const float ssq = (float)(2.0 * s);
  const float inv = float_sqrt ? (float)(1. / sqrtf(ssq)) : (float)(1. / (double)(float)sqrt((double)ssq));
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < n; e += (long long)gridDim.x * 256) {
    cf v = c[e];
    v.x *= inv;
    v.y *= inv;
    c[e] = v;
  }
}

// padded rows (mx + 2) -> dense image
__global__ void k_unpad(const float* __restrict__ src, int mx, int my, float* __restrict__ dst) {
  const long long n = (long long)mx * my;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < n; e += (long long)gridDim.x * 256)
    dst[e] = src[(e / mx) * (mx + 2) + e % mx];
}

// SwapRealSpaceQuadrants of a 2-D half spectrum: the phase of a shift by (W/2, H/2) with the
// reference's float phase arithmetic (image.cpp:9221-9268, functions.h:494-504)
__global__ void k_swap_phase_2d(cf* __restrict__ c, int W, int H) {
#pragma clang fp contract(off)
  const double PI = 3.14159265359;
  const int nh = W / 2 + 1;
  const float sx = (float)(W / 2), sy = (float)(H / 2);
  const long long total = (long long)nh * H;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
    const int i = (int)(e % nh), j = (int)(e / nh);
    const float px = sx * (float)i * 2.0 * PI / (float)W;
    const float py = sy * (float)(j >= H / 2 ? j - H : j) * 2.0 * PI / (float)H;
    const float ph = -px - py - 0.0f;
    float sn, cs;
    sincosf(ph, &sn, &cs);
    const cf v = c[e];
    c[e] = cmake(v.x * cs - v.y * sn, v.x * sn + v.y * cs);
  }
}

// the spectrum (H rows of W/2 + 1) transposed to [kx][ky], each element times the row's phase
// (pc + i ps, in double; pc = NULL for none)
__global__ void k_transpose_phase(const cf* __restrict__ spec, int Wh, int H, const double* __restrict__ pc,
                                  const double* __restrict__ ps, cf* __restrict__ out) {
  const long long total = (long long)Wh * H;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
    const int y = (int)(e % H), x = (int)(e / H);
    const cf v = spec[(long long)y * Wh + x];
    if (pc)
      out[e] = cmake((float)(v.x * pc[y] - v.y * ps[y]), (float)(v.x * ps[y] + v.y * pc[y]));
    else
      out[e] = v;
  }
}

}  // namespace
}  // namespace tmp
