// Local refinement of template-matching peaks on the GPU and its C ABI (include/tm_refine.h):
// the per-image work of refine_template (src/programs/refine_template/refine_template.cpp).
//
// Each peak's coordinate descent (refine_template.cpp:746-1086) is a C++20 coroutine that suspends
// at every TemplateScore call. The driver keeps up to max_in_flight peaks in flight, collects their
// pending TemplateScore requests into one batch, runs the batch as a few device launches (slice
// extraction, the slice's C2R and the projection's R2C, the CTF x whitening product fused with the
// spectral energy and the conjugate product with the particle window, the correlation's C2R and a
// per-item argmax), and resumes every peak with its result. Items of a batch share nothing, so a
// peak's result does not depend on which peaks were in flight with it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cfloat>
#include <climits>
#include <cmath>
#include <coroutine>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/tm_refine.h"
#include "tm_host.hpp"
#include "tm_prep_kernels.hpp"

namespace rtk {

using tmk::cf;
using tmk::cmake;
using tmk::cmul;
using tmk::cconj;

// CTF constants of CTF::Init (ctf.cpp:93-150) in pixel units.
struct CtfConst {
  float wavelength, squared_wavelength, spherical_aberration, phase_shift, amplitude_term, astigmatism_azimuth;
};

// Whitening curve lookup, Curve::ReturnLinearInterpolationFromX (curve.cpp:377-396) with the bin of
// ReturnIndexOfNearestPreviousBin (curve.cpp:834-870): the c with x[c] <= wx < x[c+1].
__device__ float curve_value(const float* __restrict__ cx, const float* __restrict__ cy, int n, float wx) {
#pragma clang fp contract(off)
  int c;
  if (wx < cx[0]) {
    c = 0;
  } else {
    c = (int)((wx - cx[0]) / (cx[1] - cx[0]));
    c = c < 0 ? 0 : (c > n - 1 ? n - 1 : c);
    while (c > 0 && wx < cx[c]) --c;
    while (c < n - 1 && wx >= cx[c + 1]) ++c;
  }
  float r = 0.0f;
  if (c == n - 1) {
    r = cy[n - 1];
  } else {
    const float d = (wx - cx[c]) / (cx[c + 1] - cx[c]);
    r = (float)((double)r + (double)cy[c] * (1.0 - (double)d));
    r += cy[c + 1] * d;
  }
  return r;
}

// Projection filter value at pixel (i, j) of a t x (t/2+1) spectrum: Image::CalculateCTFImage
// (image.cpp:3113-3165, CTF::Evaluate ctf.cpp:445-461 with ctf.cpp:497-500 and DefocusGivenAzimuth
// ctf.cpp:538-541) times Image::ApplyCurveFilter (image.cpp:7284-7320), in the reference's float and
// double steps, without contraction into fused multiply-adds.
__device__ float filter_value(int i, int j, int t, float d1, float d2, const CtfConst& k, const float* __restrict__ cx,
                              const float* __restrict__ cy, int n_curve) {
#pragma clang fp contract(off)
  const double fv = 1.0 / (double)t;
  const int ly = j >= t / 2 ? j - t : j;
  const float y2 = (float)(ly * fv);
  const float x2 = (float)(i * fv);
  const float az = (i == 0 && j == 0) ? 0.0f : atan2f(y2, x2);
  const float f2 = x2 * x2 + y2 * y2;
  float ctf;
  if (d1 == 0.0f && d2 == 0.0f) {
    ctf = -0.7f;
  } else {
    const float df = 0.5f * (d1 + d2 + cosf(2.0f * (az - k.astigmatism_azimuth)) * (d1 - d2));
    const float chi = 3.14159265359f * k.wavelength * f2 * (df - 0.5f * k.squared_wavelength * f2 * k.spherical_aberration) +
                      k.phase_shift + k.amplitude_term;
    ctf = -sinf(chi);
  }
  const double dx = i * fv, dy = ly * fv;
  const float sx = (float)(dx * dx), sy = (float)(dy * dy);
  const float s2 = sx + sy + 0.0f;
  const float sf = (float)sqrt((double)s2);
  if (!(sf <= 1.0f)) return 0.0f;
  return ctf * curve_value(cx, cy, n_curve, sf);
}

// Particle windows (refine_template.cpp:713-718): window p is the t x t centre clip of the
// whitened image after RealSpaceIntegerShift (image.cpp:9064-9107), i.e. pixel (ii, jj) reads the
// image at linear index (k - shift0[p]) mod n with k the index of (mx/2 + ii - t/2, my/2 + jj - t/2).
// Output rows of t + 2 floats (padded layout).
__global__ void k_window_gather(const float* __restrict__ img, int mx, int my, int t,
                                const long long* __restrict__ shift0, float* __restrict__ out) {
  const int jj = blockIdx.x, p = blockIdx.y;
  const long long n = (long long)mx * my;
  const int j = my / 2 + jj - t / 2;
  float* o = out + ((size_t)p * t + jj) * (t + 2);
  for (int ii = threadIdx.x; ii < t + 2; ii += blockDim.x) {
    float v = 0.0f;
    const int i = mx / 2 + ii - t / 2;
    if (ii < t && i >= 0 && i < mx && j >= 0 && j < my) {
      long long k = (long long)j * mx + i - shift0[p];
      if (k < 0) k += n;
      v = img[k];
    }
    o[ii] = v;
  }
}

// Particle mask of each window (Image::CosineMask, image.cpp:3934-4209, as refine_template.cpp:716
// calls it): one block per window. w[e] for pixel e of the t x t box is the cosine edge weight in
// [0, 1] inside the edge band, 2 inside the mask radius (kept) and -1 outside (set to the mean).
// The band mean is a double sum over the band pixels, as the reference's; band pixels become
// v e + (1 - e) mean in double, rounded to float once.
template <int NT>
__global__ __launch_bounds__(NT) void k_window_mask(float* __restrict__ win_real, int t, const float* __restrict__ w,
                                                    int n_band) {
#pragma clang fp contract(off)
  __shared__ double red[NT];
  float* o = win_real + (size_t)blockIdx.x * t * (t + 2);
  const int n = t * t;
  double s = 0.0;
  for (int e = threadIdx.x; e < n; e += NT) {
    const float m = w[e];
    if (m >= 0.0f && m <= 1.0f) s += (double)o[(e / t) * (t + 2) + e % t];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = NT / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  const double mean = red[0] / (double)n_band;
  for (int e = threadIdx.x; e < n; e += NT) {
    const float m = w[e];
    float* v = o + (e / t) * (t + 2) + e % t;
    if (m < 0.0f)
      *v = (float)mean;
    else if (m <= 1.0f)
      *v = (float)((double)(*v * m) + (1.0 - (double)m) * mean);
  }
}

// Window spectrum times the quadrant-swap phase (SwapRealSpaceQuadrants), scattered to its slot.
__global__ void k_window_phase(const cf* __restrict__ stage, const cf* __restrict__ phase, int per,
                               const int* __restrict__ slots, cf* __restrict__ win) {
  const int p = blockIdx.y;
  const cf* s = stage + (size_t)p * per;
  cf* d = win + (size_t)slots[p] * per;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < per; e += gridDim.x * blockDim.x) d[e] = cmul(s[e], phase[e]);
}

// TemplateScore (refine_template.cpp:54-72) on the projection spectra of a batch, item b:
// c = C x filter (MultiplyPixelWise), DC 0 (ZeroCentralPixel), the Hermitian-weighted sum of |c|^2
// of ReturnSumOfSquares' Fourier branch (image.cpp:185-311) into one partial per block, and
// C <- conj(c) x window. The division by sqrt(sum) is applied to the peak value afterwards (a
// positive scale does not move the peak).
template <int NT>
__global__ __launch_bounds__(NT) void k_score_product(cf* __restrict__ C, int t, const cf* __restrict__ win,
                                                      const int* __restrict__ slots, const float* __restrict__ dfs,
                                                      CtfConst k, const float* __restrict__ cx,
                                                      const float* __restrict__ cy, int n_curve,
                                                      double* __restrict__ partial) {
  __shared__ double red[NT];
  const int th = t / 2 + 1, per = t * th, b = blockIdx.y;
  cf* c = C + (size_t)b * per;
  const cf* w = win + (size_t)slots[b] * per;
  const float d1 = dfs[2 * b], d2 = dfs[2 * b + 1];
  double sum = 0.0;
  for (int e = blockIdx.x * NT + threadIdx.x; e < per; e += gridDim.x * NT) {
    const int j = e / th, i = e - j * th;
    const float f = filter_value(i, j, t, d1, d2, k, cx, cy, n_curve);
    cf v = c[e];
    v = cmake(v.x * f, v.y * f);
    if (e == 0) v = cmake(0.0f, 0.0f);
    const int jj = j >= t / 2 ? j - t : j;
    const bool edge_x = (i == 0 || i == t / 2);
    const double a = (double)hypotf(v.x, v.y);
    const double a2 = a * a;
    if (edge_x && (jj == 0 || jj == -t / 2)) sum += a2 * 0.5;
    else if (!edge_x || jj >= 0) sum += a2;
    c[e] = cmul(cconj(v), w[e]);
  }
  red[threadIdx.x] = sum;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(size_t)b * gridDim.x + blockIdx.x] = red[0];
}

// Peak of each correlation image (Image::FindPeakWithIntegerCoordinates with no radius or edge
// limit, image.cpp:9961-10105): the first maximum in scan order. Stage 1: block q of item b scans
// rows [q t / nq, (q + 1) t / nq) and writes its (value, index) partial.
template <int NT>
__global__ __launch_bounds__(NT) void k_score_peak_rows(const float* __restrict__ R, int t, float* __restrict__ pv,
                                                        int* __restrict__ pi) {
  __shared__ float rv[NT];
  __shared__ int ri[NT];
  const int q = blockIdx.x, nq = gridDim.x, b = blockIdx.y;
  const float* r = R + (size_t)b * t * (t + 2);
  const int e0 = q * t / nq * t, e1 = (q + 1) * t / nq * t;
  float best = -FLT_MAX;
  int bi = 0x7fffffff;
  for (int e = e0 + threadIdx.x; e < e1; e += NT) {
    const int j = e / t, i = e - j * t;
    const float v = r[(size_t)j * (t + 2) + i];
    if (v > best) { best = v; bi = e; }
  }
  rv[threadIdx.x] = best;
  ri[threadIdx.x] = bi;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const float v = rv[threadIdx.x + s];
      const int k = ri[threadIdx.x + s];
      if (v > rv[threadIdx.x] || (v == rv[threadIdx.x] && k < ri[threadIdx.x])) { rv[threadIdx.x] = v; ri[threadIdx.x] = k; }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[(size_t)b * nq + q] = rv[0];
    pi[(size_t)b * nq + q] = ri[0];
  }
}

// Stage 2, one thread per item: the partials in row order, the peak reported relative to the box
// centre, its value divided by sqrtf(ReturnSumOfSquares) of the filtered projection (the division
// TemplateScore applies to the spectrum, refine_template.cpp:60).
__global__ void k_score_peak_final(const float* __restrict__ pv, const int* __restrict__ pi, int nq, int t,
                                   const double* __restrict__ partial, int nparts, int n, float* __restrict__ out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  float best = -FLT_MAX;
  int bi = 0x7fffffff;
  for (int q = 0; q < nq; ++q) {
    const float v = pv[(size_t)b * nq + q];
    const int k = pi[(size_t)b * nq + q];
    if (v > best || (v == best && k < bi)) { best = v; bi = k; }
  }
  double ss = 0.0;
  for (int q = 0; q < nparts; ++q) ss += partial[(size_t)b * nparts + q];
  const float sos = (float)(2.0 * ss);
  const float inv = (float)(1. / sqrtf(sos));
  const int e = bi == 0x7fffffff ? 0 : bi;
  out[3 * b] = (float)(e % t - t / 2);
  out[3 * b + 1] = (float)(e / t - t / 2);
  out[3 * b + 2] = bi == 0x7fffffff ? -FLT_MAX : best * inv;
}

// Peak list on the scaled MIP (refine_template.cpp:561-601): first maximum in scan order over the
// pixels at least `edge` from the borders (upper bounds inclusive), as (value, dense index) partials
// per block, then one block combines them.
template <int NT>
__global__ __launch_bounds__(NT) void k_map_argmax(const float* __restrict__ a, int mx, int my, int edge,
                                                   float* __restrict__ pv, long long* __restrict__ pi) {
  __shared__ float rv[NT];
  __shared__ long long ri[NT];
  const long long n = (long long)mx * my;
  float best = -FLT_MAX;
  long long bi = LLONG_MAX;
  for (long long e = (long long)blockIdx.x * NT + threadIdx.x; e < n; e += (long long)gridDim.x * NT) {
    const int j = (int)(e / mx), i = (int)(e - (long long)j * mx);
    if (i >= edge && i <= mx - edge && j >= edge && j <= my - edge) {
      const float v = a[e];
      if (v > best) { best = v; bi = e; }
    }
  }
  rv[threadIdx.x] = best;
  ri[threadIdx.x] = bi;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const float v = rv[threadIdx.x + s];
      const long long q = ri[threadIdx.x + s];
      if (v > rv[threadIdx.x] || (v == rv[threadIdx.x] && q < ri[threadIdx.x])) { rv[threadIdx.x] = v; ri[threadIdx.x] = q; }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { pv[blockIdx.x] = rv[0]; pi[blockIdx.x] = ri[0]; }
}

template <int NT>
__global__ __launch_bounds__(NT) void k_argmax_final(const float* __restrict__ pv, const long long* __restrict__ pi, int n,
                                                     float* __restrict__ ov, long long* __restrict__ oi) {
  __shared__ float rv[NT];
  __shared__ long long ri[NT];
  float best = -FLT_MAX;
  long long bi = LLONG_MAX;
  for (int e = threadIdx.x; e < n; e += NT) {
    const float v = pv[e];
    const long long q = pi[e];
    if (v > best || (v == best && q < bi)) { best = v; bi = q; }
  }
  rv[threadIdx.x] = best;
  ri[threadIdx.x] = bi;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const float v = rv[threadIdx.x + s];
      const long long q = ri[threadIdx.x + s];
      if (v > rv[threadIdx.x] || (v == rv[threadIdx.x] && q < ri[threadIdx.x])) { rv[threadIdx.x] = v; ri[threadIdx.x] = q; }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { *ov = rv[0]; *oi = ri[0]; }
}

// Pixels whose squared distance from (cx, cy) is <= radius become -FLT_MAX (refine_template.cpp:580-592;
// the program compares the squared distance with min_peak_radius). Launched over a box that holds
// every such pixel.
__global__ void k_map_mask(float* __restrict__ a, int mx, int x0, int y0, int w, int hgt, float cx, float cy,
                           float radius) {
#pragma clang fp contract(off)
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= w * hgt) return;
  const int j = y0 + e / w, i = x0 + e % w;
  const float dy = (float)j - cy, dx = (float)i - cx;
  const float sq = (float)((double)dx * dx) + (float)((double)dy * dy);
  if (sq <= radius) a[(size_t)j * mx + i] = -FLT_MAX;
}


// One TemplateScore per workgroup with every intermediate in LDS, for t = P (no padding) and no
// pixel-size change: the steps of the multi-kernel path (k_extract_cols<t, 8, 1>, k_rows_c2r<t, 64>,
// k_rows_r2c<t, 64> scaled, k_cols<t, 64, -1, 0>, k_score_product<256>, k_cols<t, 64, +1, 0>,
// k_rows_c2r<t, 64>, k_score_peak_rows<256> + k_score_peak_final) on the same per-line FFT plans
// (16 threads per line, 16 lines at a time: a butterfly's arithmetic does not depend on which thread
// runs it) and the same 256-thread reductions (one part each), so the results are bit-identical to
// that path. Lines: col[i] is column i of the t x (t/2+1) spectrum,
// row[j] row j, re[j] real row j (in col's memory).
template <int T>
__global__ __launch_bounds__(256) void k_score_fused(const cf* __restrict__ tpl, const float* __restrict__ mats,
                                                     const cf* __restrict__ phase, const cf* __restrict__ tw_t,
                                                     const cf* __restrict__ tw_h, const cf* __restrict__ win,
                                                     const int* __restrict__ slots, const float* __restrict__ dfs,
                                                     CtfConst kc, const float* __restrict__ cx,
                                                     const float* __restrict__ cy, int n_curve, float scale,
                                                     float* __restrict__ out) {
  constexpr int NT = 256, TH = T / 2 + 1, H = T / 2, LT = 16, LINES = NT / LT;
  constexpr int NCOL = (TH + LINES - 1) / LINES * LINES;
  // re shares col's memory: the real rows are written only after col has been copied into row, and
  // read back into row before col is written again
  __shared__ cf col[NCOL][tmk::lds_len(T)];
  __shared__ cf row[T][tmk::lds_len(H + 1)];
  static_assert(sizeof(float) * T * (T + 2) <= sizeof(cf) * NCOL * tmk::lds_len(T), "re must fit in col");
  float(*re)[T + 2] = reinterpret_cast<float(*)[T + 2]>(&col[0][0]);
  __shared__ double red[NT];
  __shared__ float rv[NT];
  __shared__ int ri[NT];
  const int b = blockIdx.x, tid = threadIdx.x, lt = tid % LT, ln = tid / LT;
  const float* m = mats + (size_t)b * 9;
  // slice and its inverse column transforms
  for (int e = tid; e < NCOL * T; e += NT) {
    const int c = e % NCOL, j = e / NCOL;
    col[c][tmk::lpad(j)] = c < TH ? tmk::slice_value<1>(tpl, T, m, c, j, phase, nullptr) : cmake(0.f, 0.f);
  }
  __syncthreads();
  for (int c0 = 0; c0 < NCOL; c0 += LINES) tmk::fft_inplace<T, +1, LT>(col[c0 + ln], tw_t, lt);
  // rows from columns, then C2R of each row into re (k_rows_c2r)
  auto cols_to_rows = [&]() {
    __syncthreads();
    for (int e = tid; e < T * TH; e += NT) {
      const int j = e / TH, i = e - j * TH;
      row[j][tmk::lpad(i)] = col[i][tmk::lpad(j)];
    }
    __syncthreads();
    for (int r0 = 0; r0 < T; r0 += LINES) {
      cf* buf = row[r0 + ln];
      for (int k = lt; k <= H / 2; k += LT) {
        if (k == 0) {
          const float x0 = buf[tmk::lpad(0)].x, xh = buf[tmk::lpad(H)].x;
          buf[tmk::lpad(0)] = cmake(x0 + xh, x0 - xh);
        } else {
          const int kk = H - k;
          const cf Xk = buf[tmk::lpad(k)], Xc = cconj(buf[tmk::lpad(kk)]);
          const cf E = tmk::cadd(Xk, Xc);
          const cf D = tmk::csub(Xk, Xc);
          const cf w = tw_t[k];
          const cf O = cmul(D, cconj(w));
          const cf Ok = cmul(cconj(D), w);
          buf[tmk::lpad(k)] = cmake(E.x - O.y, E.y + O.x);
          buf[tmk::lpad(kk)] = cmake(E.x - Ok.y, -E.y + Ok.x);
        }
      }
      __syncthreads();
      tmk::fft_inplace<H, +1, LT>(buf, tw_h, lt);
      for (int q = lt; q < H; q += LT) {
        const cf z = buf[tmk::lpad(q)];
        re[r0 + ln][2 * q] = z.x;
        re[r0 + ln][2 * q + 1] = z.y;
      }
    }
    __syncthreads();
  };
  cols_to_rows();
  // forward R2C of each row, scaled (k_rows_r2c), then the forward column transforms
  for (int r0 = 0; r0 < T; r0 += LINES) {
    cf* buf = row[r0 + ln];
    for (int q = lt; q < H; q += LT) buf[tmk::lpad(q)] = cmake(re[r0 + ln][2 * q], re[r0 + ln][2 * q + 1]);
    __syncthreads();
    tmk::fft_inplace<H, -1, LT>(buf, tw_h, lt);
    for (int k = lt; k <= H / 2; k += LT) {
      if (k == 0) {
        const cf z0 = buf[tmk::lpad(0)];
        buf[tmk::lpad(0)] = cmake((z0.x + z0.y) * scale, 0.f);
        buf[tmk::lpad(H)] = cmake((z0.x - z0.y) * scale, 0.f);
      } else {
        const int kk = H - k;
        const cf zk = buf[tmk::lpad(k)], zc = cconj(buf[tmk::lpad(kk)]);
        const cf E = tmk::cscale(tmk::cadd(zk, zc), 0.5f);
        const cf D = tmk::cscale(tmk::csub(zk, zc), 0.5f);
        const cf O = cmake(D.y, -D.x);
        const cf wO = cmul(tw_t[k], O);
        const cf Xk = tmk::cadd(E, wO);
        const cf Xkk = cconj(tmk::csub(E, wO));
        buf[tmk::lpad(k)] = tmk::cscale(Xk, scale);
        buf[tmk::lpad(kk)] = tmk::cscale(Xkk, scale);
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < T * TH; e += NT) {
    const int i = e / T, j = e - i * T;
    col[i][tmk::lpad(j)] = row[j][tmk::lpad(i)];
  }
  __syncthreads();
  for (int c0 = 0; c0 < NCOL; c0 += LINES) tmk::fft_inplace<T, -1, LT>(col[c0 + ln], tw_t, lt);
  __syncthreads();
  // filter, sum of squares and conj(projection) x window (k_score_product, one part)
  const cf* w = win + (size_t)slots[b] * T * TH;
  const float d1 = dfs[2 * b], d2 = dfs[2 * b + 1];
  double sum = 0.0;
  for (int e = tid; e < T * TH; e += NT) {
    const int j = e / TH, i = e - j * TH;
    const float f = filter_value(i, j, T, d1, d2, kc, cx, cy, n_curve);
    cf v = col[i][tmk::lpad(j)];
    v = cmake(v.x * f, v.y * f);
    if (e == 0) v = cmake(0.0f, 0.0f);
    const int jj = j >= T / 2 ? j - T : j;
    const bool edge_x = (i == 0 || i == T / 2);
    const double a = (double)hypotf(v.x, v.y);
    const double a2 = a * a;
    if (edge_x && (jj == 0 || jj == -T / 2)) sum += a2 * 0.5;
    else if (!edge_x || jj >= 0) sum += a2;
    col[i][tmk::lpad(j)] = cmul(cconj(v), w[e]);
  }
  red[tid] = sum;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  // inverse column transforms, rows, C2R
  for (int c0 = 0; c0 < NCOL; c0 += LINES) tmk::fft_inplace<T, +1, LT>(col[c0 + ln], tw_t, lt);
  cols_to_rows();
  // first maximum in scan order (k_score_peak_rows, one part) and the result (k_score_peak_final)
  float best = -FLT_MAX;
  int bi = 0x7fffffff;
  for (int e = tid; e < T * T; e += NT) {
    const int j = e / T, i = e - j * T;
    const float v = re[j][i];
    if (v > best) { best = v; bi = e; }
  }
  rv[tid] = best;
  ri[tid] = bi;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (tid < s) {
      const float v = rv[tid + s];
      const int k = ri[tid + s];
      if (v > rv[tid] || (v == rv[tid] && k < ri[tid])) { rv[tid] = v; ri[tid] = k; }
    }
    __syncthreads();
  }
  if (tid == 0) {
    float pbest = -FLT_MAX;
    int pbi = 0x7fffffff;
    if (rv[0] > pbest || (rv[0] == pbest && ri[0] < pbi)) { pbest = rv[0]; pbi = ri[0]; }
    double ss = 0.0;
    ss += red[0];
    const float sos = (float)(2.0 * ss);
    const float inv = (float)(1. / sqrtf(sos));
    const int e = pbi == 0x7fffffff ? 0 : pbi;
    out[3 * b] = (float)(e % T - T / 2);
    out[3 * b + 1] = (float)(e / T - T / 2);
    out[3 * b + 2] = pbi == 0x7fffffff ? -FLT_MAX : pbest * inv;
  }
}

}  // namespace rtk

namespace {

using tmk::cf;
thread_local std::string g_rerr;

// ------------------------------------------------------------------ descent coroutine
struct ScoreRequest {
  float phi, theta, psi;  // degrees
  float df1, df2;         // SetDefocus arguments (pixels)
  float factor;           // TemplateScore's pixel_size_factor
};
struct ScoreResult {
  float x, y, value;
};

struct DescentTask {
  struct promise_type {
    ScoreRequest req{};
    ScoreResult res{};
    std::exception_ptr error;
    DescentTask get_return_object() { return DescentTask{std::coroutine_handle<promise_type>::from_promise(*this)}; }
    std::suspend_always initial_suspend() noexcept { return {}; }
    std::suspend_always final_suspend() noexcept { return {}; }
    void return_void() noexcept {}
    void unhandled_exception() noexcept { error = std::current_exception(); }
  };
  std::coroutine_handle<promise_type> h;
  explicit DescentTask(std::coroutine_handle<promise_type> c) : h(c) {}
  DescentTask(DescentTask&& o) noexcept : h(std::exchange(o.h, {})) {}
  DescentTask& operator=(DescentTask&& o) noexcept {
    if (this != &o) {
      if (h) h.destroy();
      h = std::exchange(o.h, {});
    }
    return *this;
  }
  DescentTask(const DescentTask&) = delete;
  ~DescentTask() {
    if (h) h.destroy();
  }
};

// co_await Score{...} suspends the peak until the batch holding its request has run.
struct Score {
  ScoreRequest r;
  std::coroutine_handle<DescentTask::promise_type> c{};
  bool await_ready() const noexcept { return false; }
  void await_suspend(std::coroutine_handle<DescentTask::promise_type> cc) noexcept {
    c = cc;
    c.promise().req = r;
  }
  ScoreResult await_resume() const noexcept { return c.promise().res; }
};

// The per-thread result maps best_*_local of the program (refine_template.cpp:680-689) as the input
// maps plus the addresses written so far. Addresses are those of the padded layout (rows of mx + 2);
// the two padding columns read as 0.
enum { M_PSI, M_THETA, M_PHI, M_DEFOCUS, M_PIXEL_SIZE };
struct LocalMaps {
  const float* in[5] = {};  // psi, theta, phi, defocus, pixel_size (dense mx * my)
  int mx = 0, my = 0;
  std::unordered_map<long long, std::array<float, 5>> written;
  float base(long long addr, int f) const {
    const long long rs = mx + 2;
    if (addr < 0 || addr >= rs * my) throw TmError{TM_ERR_INVALID, "refined peak address outside the maps"};
    const long long row = addr / rs, col = addr % rs;
    return col < mx ? in[f][row * mx + col] : 0.0f;
  }
  float get(long long addr, int f) const {
    auto it = written.find(addr);
    return it != written.end() ? it->second[f] : base(addr, f);
  }
  void set(long long addr, int f, float v) {
    auto it = written.find(addr);
    if (it == written.end()) {
      std::array<float, 5> a;
      for (int q = 0; q < 5; ++q) a[q] = base(addr, q);
      it = written.emplace(addr, a).first;
    }
    it->second[f] = v;
  }
};

struct RunConst {
  tm_refine_params p;
  const float* mip = nullptr;         // dense
  const float* scaled_mip = nullptr;  // dense
  float defocus_search_range = 0, defocus_refine_step = 0, defocus_step = 0, sqrt_t = 0;
  float pixel_size_search_range = 0, pixel_size_refine_step = 0;
  long long rs = 0;  // padded row stride mx + 2
};

struct PeakState {
  float x = 0, y = 0, value = 0;  // found peak, relative to the image centre
  LocalMaps* maps = nullptr;
  tm_refined_peak rec{};
  int evaluations = 0;
};

inline int myroundint(float a) { return a > 0 ? int(a + 0.5) : int(a - 0.5); }

float dense_at(const float* m, int mx, long long rs, long long addr) {
  const long long row = addr / rs, col = addr % rs;
  return col < mx ? m[row * mx + col] : 0.0f;
}

// The program's per-peak body (refine_template.cpp:704-1113 with pixel_size_search_range 0).
DescentTask refine_peak(const RunConst& K, PeakState& st) {
  const tm_refine_params& p = K.p;
  LocalMaps& L = *st.maps;
  const float box_cx = (float)(p.image_x / 2), box_cy = (float)(p.image_y / 2);
  const float px = st.x + box_cx, py = st.y + box_cy;
  const long long current_address = (long long)myroundint(py) * K.rs + myroundint(px);
  float current_phi = L.get(current_address, M_PHI);
  float current_theta = L.get(current_address, M_THETA);
  float current_psi = L.get(current_address, M_PSI);
  float current_defocus = L.get(current_address, M_DEFOCUS);
  float current_pixel_size = L.get(current_address, M_PIXEL_SIZE);
  const float scaled_here = dense_at(K.scaled_mip, p.image_x, K.rs, current_address);
  const float mip_here = dense_at(K.mip, p.image_x, K.rs, current_address);
  float best_score = -FLT_MAX, score;
  long long best_address = current_address;
  ScoreResult best_peak{0, 0, 0}, tp{0, 0, 0};
  float offset_distance = 0.f;
  const float score_adjustment = 1.0f;
  float factor = 1.0f;  // template_object.pixel_size_factor (refine_template.cpp:720)
  auto req = [&](float cur, float step, float phi, float theta, float psi) {
    ++st.evaluations;
    return ScoreRequest{phi,    theta, psi, (p.defocus1 + cur + step) / p.pixel_size, (p.defocus2 + cur + step) / p.pixel_size,
                        factor};
  };
  auto improve = [&](const ScoreResult& r) {
    best_peak = r;
    best_score = r.value;
    const long long address_offset = K.rs * myroundint(r.y) + myroundint(r.x);
    best_address = current_address + address_offset;
    offset_distance = sqrtf(powf(r.x, 2) + powf(r.y, 2));
  };
  tp = co_await Score{req(current_defocus, 0.0f, current_phi, current_theta, current_psi)};
  float starting_score = tp.value * K.sqrt_t;
  starting_score = score_adjustment * scaled_here * starting_score / mip_here;
  if (K.defocus_search_range != 0.0f) {
    const int nd = myroundint(float(K.defocus_search_range) / float(K.defocus_step));
    for (int defocus_is = -nd; defocus_is <= nd; ++defocus_is) {
      tp = co_await Score{req(current_defocus, defocus_is * K.defocus_step, current_phi, current_theta, current_psi)};
      score = tp.value;
      if (score > best_score) {
        improve(tp);
        L.set(best_address, M_DEFOCUS, current_defocus + defocus_is * K.defocus_step);
      }
    }
    current_defocus = L.get(best_address, M_DEFOCUS);
  }
  int phi_i = 0, theta_i = 0, psi_i = 0, defocus_i = 0, size_i = 0;
  // block 0: defocus refinement with nested angle descents (:848-929) at pixel_size_factor 1;
  // block 1: the pixel-size block (:986-1086), after restarting from the values at best_address and
  // the pixel-size scan (:931-984)
  for (int block = 0; block < 2; ++block) {
    float filter_step = 0.0f;
    const float mult_start = block == 0 ? K.defocus_step / K.defocus_refine_step : 1.0f;
    const float outer_step = block == 0 ? K.defocus_refine_step : K.pixel_size_refine_step;
    for (float mult_i = mult_start; mult_i > 0.5f; mult_i /= 2.0f) {
      for (int ll = 0; ll < 2; ll = -2 * ll + 1) {
        if ((ll != 0) && (outer_step == 0.0f)) break;
        float best_outer_score;
        do {
          best_outer_score = best_score;
          float written_defocus = current_defocus, written_pixel_size = current_pixel_size;
          if (block == 0) {
            if (K.defocus_search_range != 0.0f) defocus_i += myroundint(mult_i * ll);
            filter_step = defocus_i * K.defocus_refine_step;
            written_defocus = current_defocus + defocus_i * K.defocus_refine_step;
          } else {
            if (K.pixel_size_search_range != 0.0f) size_i += myroundint(mult_i * ll);
            factor = (p.pixel_size + current_pixel_size + float(size_i) * K.pixel_size_refine_step) / p.pixel_size;
            written_pixel_size = current_pixel_size + size_i * K.pixel_size_refine_step;
          }
          for (int kk = 0; kk < 2; kk = -2 * kk + 1) {
            float best_phi_score;
            do {
              best_phi_score = best_score;
              phi_i += kk;
              for (int jj = 0; jj < 2; jj = -2 * jj + 1) {
                float best_theta_score;
                do {
                  best_theta_score = best_score;
                  theta_i += jj;
                  for (int ii = 0; ii < 2; ii = -2 * ii + 1) {
                    float best_psi_score;
                    do {
                      best_psi_score = best_score;
                      psi_i += ii;
                      tp = co_await Score{req(current_defocus, filter_step, current_phi + phi_i * p.angular_step,
                                              current_theta + theta_i * p.angular_step,
                                              current_psi + psi_i * p.in_plane_angular_step)};
                      score = tp.value;
                      if (score > best_score) {
                        improve(tp);
                        L.set(best_address, M_PSI, current_psi + psi_i * p.in_plane_angular_step);
                        L.set(best_address, M_THETA, current_theta + theta_i * p.angular_step);
                        L.set(best_address, M_PHI, current_phi + phi_i * p.angular_step);
                        L.set(best_address, M_DEFOCUS, written_defocus);
                        L.set(best_address, M_PIXEL_SIZE, written_pixel_size);
                      }
                    } while (best_score > best_psi_score);
                    psi_i -= ii;
                  }
                } while (best_score > best_theta_score);
                theta_i -= jj;
              }
            } while (best_score > best_phi_score);
            phi_i -= kk;
          }
        } while (best_score > best_outer_score);
        if (block == 0 && K.defocus_search_range != 0.0f) defocus_i -= ll;
        if (block == 1 && K.pixel_size_search_range != 0.0f) size_i -= ll;
      }
    }
    if (block == 0) {
      current_phi = L.get(best_address, M_PHI);
      current_theta = L.get(best_address, M_THETA);
      current_psi = L.get(best_address, M_PSI);
      current_defocus = L.get(best_address, M_DEFOCUS);
      current_pixel_size = L.get(best_address, M_PIXEL_SIZE);
      phi_i = theta_i = psi_i = 0;
      filter_step = 0.0f;
      // the pixel-size scan at the restart values (:946-984)
      if (K.pixel_size_search_range != 0.0f) {
        const int ns = myroundint(float(K.pixel_size_search_range) / float(K.pixel_size_refine_step));
        for (int size_is = -ns; size_is <= ns; ++size_is) {
          factor = (p.pixel_size + current_pixel_size + float(size_is) * K.pixel_size_refine_step) / p.pixel_size;
          tp = co_await Score{req(current_defocus, 0.0f, current_phi, current_theta, current_psi)};
          score = tp.value;
          if (score > best_score) {
            improve(tp);
            L.set(best_address, M_PIXEL_SIZE, current_pixel_size + size_is * K.pixel_size_refine_step);
          }
        }
        current_pixel_size = L.get(best_address, M_PIXEL_SIZE);
      }
      phi_i = theta_i = psi_i = 0;
      size_i = 0;
    }
  }
  tm_refined_peak& r = st.rec;
  r.x = st.x;
  r.y = st.y;
  r.value = st.value;
  r.psi = L.base(current_address, M_PSI);
  r.theta = L.base(current_address, M_THETA);
  r.phi = L.base(current_address, M_PHI);
  r.defocus = L.base(current_address, M_DEFOCUS);
  r.pixel_size = L.base(current_address, M_PIXEL_SIZE);
  r.x_pos = st.x + best_peak.x;
  r.y_pos = st.y + best_peak.y;
  r.shift_x = best_peak.x;
  r.shift_y = best_peak.y;
  r.r_psi = L.get(best_address, M_PSI);
  r.r_theta = L.get(best_address, M_THETA);
  r.r_phi = L.get(best_address, M_PHI);
  r.r_defocus = L.get(best_address, M_DEFOCUS);
  r.r_pixel_size = L.get(best_address, M_PIXEL_SIZE);
  r.peak_height = score_adjustment * scaled_here * best_score / mip_here * K.sqrt_t;
  r.mip = score_adjustment * best_score * K.sqrt_t;
  r.best_score = best_score;
  r.starting_score = starting_score;
  r.offset = offset_distance;
  r.best_address = best_address;
  r.evaluations = st.evaluations;
}

template <class T>
struct PinnedBuf {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    TM_HIP(hipHostMalloc((void**)&p, count * sizeof(T), hipHostMallocDefault));
    n = count;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
  ~PinnedBuf() { release(); }
};

}  // namespace

struct tm_refine {
  int device = 0;
  hipStream_t stream = nullptr;
  bool initialised = false;
  tm_refine_params prm{};
  int t = 0, th = 0, P = 0, Ph = 0, K = 0;  // K: max_in_flight
  int n_curve = 0;
  rtk::CtfConst ctf{};
  DevBuf<float> img, curve_x, curve_y, mats, dfs, results, smip, win_real, arg_v;
  DevBuf<float> mask_w;  // particle-mask weights of the t x t window (empty: no mask)
  int mask_band = 0;     // pixels in the mask's edge band
  DevBuf<cf> tpl, phaseP, phaseT, slice, proj, win, win_stage;
  DevBuf<double> partial;
  DevBuf<float> peak_v;
  DevBuf<int> slots, peak_i;
  DevBuf<long long> shifts, arg_i;
  PinnedBuf<float> h_mats, h_dfs, h_results, h_arg_v;
  PinnedBuf<int> h_slots;
  PinnedBuf<long long> h_shifts, h_arg_i;
  int score_parts = 1, peak_parts = 1;
  bool fused = true;  // one-workgroup TemplateScore (k_score_fused) where it applies
  std::map<int, cf*> tw;
  // pixel-size resampling of projections (allocated at the first item that needs it)
  DevBuf<float> rs_pad, rs_big, rs_clip;
  // device preparation of a micrograph (tm_refine_set_micrograph)
  DevBuf<float> pr_src, pr_pad, pr_st;
  DevBuf<cf> pr_spec;
  DevBuf<double> pr_part, pr_acc;
  DevBuf<cf> rs_spec, rs_new;
  struct BluesteinTables {
    int M = 0;
    DevBuf<cf> chirp, bhat;
  };
  std::map<std::pair<int, int>, std::unique_ptr<BluesteinTables>> bluestein;
  // timing: event pairs around each stage of a TemplateScore batch, resolved after the batch's
  // synchronisation
  enum { S_EXTRACT, S_SLICE_C2R, S_PROJ_R2C, S_PRODUCT, S_CCF_C2R, S_PEAK, S_FUSED, S_COUNT };
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<std::pair<int, size_t>> ev_marks;
  double stage_ms[S_COUNT] = {};
  long long stage_launches[S_COUNT] = {};
  long long stage_items[S_COUNT] = {};
  int batch_items = 0;

  hipEvent_t batch_done[2] = {nullptr, nullptr};  // the fused path's two batch halves

  ~tm_refine() {
    for (auto& kv : tw) (void)hipFree(kv.second);
    for (auto& e : ev_pool) (void)hipEventDestroy(e);
    for (auto& e : batch_done)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
  hipEvent_t next_event() {
    if (ev_used == ev_pool.size()) {
      hipEvent_t e;
      TM_HIP(hipEventCreate(&e));
      ev_pool.push_back(e);
    }
    return ev_pool[ev_used++];
  }
  void resolve_timing() {
    for (auto& m : ev_marks) {
      float ms = 0.f;
      TM_HIP(hipEventElapsedTime(&ms, ev_pool[m.second], ev_pool[m.second + 1]));
      stage_ms[m.first] += ms;
      stage_launches[m.first] += 1;
      stage_items[m.first] += batch_items;
    }
    ev_marks.clear();
    ev_used = 0;
  }
  const cf* twiddle(int n) {
    auto it = tw.find(n);
    if (it != tw.end()) return it->second;
    auto host = twiddles(n);
    cf* d = nullptr;
    TM_HIP(hipMalloc((void**)&d, sizeof(cf) * n));
    TM_HIP(hipMemcpy(d, host.data(), sizeof(cf) * n, hipMemcpyHostToDevice));
    tw[n] = d;
    return d;
  }
};

namespace {

const char* kStageNames[tm_refine::S_COUNT] = {"extract_slice_cols", "slice_rows_c2r", "projection_r2c",
                                               "score_product",      "correlation_c2r", "score_peak",
                                               "score_fused"};

// Records an event pair around one stage on the handle's stream when timing is on.
struct Stage {
  tm_refine* h;
  int k;
  size_t start = 0;
  Stage(tm_refine* hh, int kk) : h(hh), k(kk) {
    if (h->timing) {
      start = h->ev_used;
      TM_HIP(hipEventRecord(h->next_event(), h->stream));
    }
  }
  ~Stage() {
    if (h->timing) {
      hipEvent_t e = nullptr;
      try { e = h->next_event(); } catch (...) { return; }
      if (hipEventRecord(e, h->stream) == hipSuccess) h->ev_marks.push_back({k, start});
    }
  }
};

template <class F>
int rguard(F&& f) {
  try {
    f();
    return TM_OK;
  } catch (const TmError& e) {
    g_rerr = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    g_rerr = e.what();
    return TM_ERR_INVALID;
  }
}

bool template_size_compiled(int n) {
  switch (n) {
#define TM_CASE(k) case k:
    TM_TEMPLATE_SIZES(TM_CASE)
#undef TM_CASE
    return true;
    default: return false;
  }
}

void require_ready(tm_refine* h) {
  TM_REQUIRE(h, TM_ERR_INVALID, "handle is NULL");
  TM_REQUIRE(h->initialised, TM_ERR_STATE, "tm_refine_init has not succeeded on this handle");
  TM_HIP(hipSetDevice(h->device));
}

// RealSpaceIntegerShift's starting index (image.cpp:9072-9074), with its float remainder.
long long shift_start(int mx, int my, int sx, int sy) {
  const long long n = (long long)mx * my;
  long long s = -(long long)sx - (long long)mx * sy;
  s = (long long)remainderf(float(s), float(n));
  if (s < 0) s += n;
  return s;
}

// Weights of the particle mask (Image::CosineMask, image.cpp:3934-4209, centred real-space branch,
// not inverted) for the t x t window, in the reference's float / double expression order: the
// cosine edge weight in the band r0^2 <= d^2 <= (r0 + edge)^2 (r0 = radius - edge / 2, at least
// 0), 2 inside (kept), -1 outside (the band mean). *n_band receives the band's pixel count.
std::vector<float> mask_weights(int t, float wanted_mask_radius, float wanted_mask_edge, int* n_band) {
  const double PI = 3.14159265359;
  float mask_radius = wanted_mask_radius - wanted_mask_edge * 0.5;
  if (mask_radius < 0.0) mask_radius = 0.0;
  const float mask_radius_plus_edge = mask_radius + wanted_mask_edge;
  const float r2 = powf(mask_radius, 2), re2 = powf(mask_radius_plus_edge, 2);
  const int c = t / 2;
  std::vector<float> w((size_t)t * t);
  int nb = 0;
  for (int j = 0; j < t; ++j) {
    const float y = powf(j - c, 2);
    for (int i = 0; i < t; ++i) {
      const float d2 = powf(i - c, 2) + y + 0.0f;
      float& o = w[(size_t)j * t + i];
      if (d2 >= r2 && d2 <= re2) {
        const float d = sqrtf(d2);
        o = (1.0 + cosf(PI * (d - mask_radius) / wanted_mask_edge)) / 2.0;
        ++nb;
      } else {
        o = d2 >= re2 ? -1.0f : 2.0f;
      }
    }
  }
  *n_band = nb;
  return w;
}

// Particle windows for the peaks at (xs[q], ys[q]) into window slots slot_of[q].
void make_windows(tm_refine* h, const std::vector<int>& xs, const std::vector<int>& ys, const std::vector<int>& slot_of) {
  const int n = (int)xs.size();
  if (n == 0) return;
  const int t = h->t, th = h->th;
  for (int q0 = 0; q0 < n; q0 += h->K) {
    const int m = std::min(h->K, n - q0);
    for (int q = 0; q < m; ++q) {
      h->h_shifts.p[q] = shift_start(h->prm.image_x, h->prm.image_y, xs[q0 + q], ys[q0 + q]);
      h->h_slots.p[q] = slot_of[q0 + q];
    }
    TM_HIP(hipMemcpyAsync(h->shifts.p, h->h_shifts.p, m * sizeof(long long), hipMemcpyHostToDevice, h->stream));
    TM_HIP(hipMemcpyAsync(h->slots.p, h->h_slots.p, m * sizeof(int), hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(rtk::k_window_gather, dim3(t, m), dim3(256), 0, h->stream, h->img.p, h->prm.image_x,
                       h->prm.image_y, t, h->shifts.p, h->win_real.p);
    check_launch("k_window_gather");
    if (h->mask_band > 0) {
      hipLaunchKernelGGL(rtk::k_window_mask<256>, dim3(m), dim3(256), 0, h->stream, h->win_real.p, t, h->mask_w.p,
                         h->mask_band);
      check_launch("k_window_mask");
    }
    // ForwardFFT (scaled 1/N) of each window
    const float scale = (float)(1. / (double)(float)((long long)t * t));
    rows_r2c(h, t, h->win_real.p, (size_t)t * (t + 2), t + 2, 0, 0, t, 0, h->win_stage.p, (size_t)t * th, th, t, m, scale);
    cols(h, t, h->win_stage.p, th, (size_t)t * th, m, -1, 0, nullptr, 1.0f);
    hipLaunchKernelGGL(rtk::k_window_phase, dim3(std::max(1, (t * th + 255) / 256), m), dim3(256), 0, h->stream,
                       h->win_stage.p, h->phaseT.p, t * th, h->slots.p, h->win.p);
    check_launch("k_window_phase");
    // the staging buffers are reused by the next chunk only after these launches (same stream)
    TM_HIP(hipStreamSynchronize(h->stream));
  }
}

// Edges ChangePixelSize works at for a square image of edge a (image.cpp:8731-8775): the
// intermediate (pad) and the new edge.
void change_pixel_size_dims(int a, float factor, float tol, int* pad, int* nw) {
  auto ev = [](int n) { return (n & 1) == 0; };
  if (tol < 1.0f / (a * factor)) {
    const float inter = 1.0f / (tol * a);
    int p = a * inter;
    if (ev(a) != ev(p)) p++;
    int q = int(p / factor);
    if (ev(a) != ev(q)) q++;
    *pad = p;
    *nw = q;
  } else {
    int p = a * factor;
    if (ev(a) != ev(p)) p++;
    *pad = p;
    *nw = a;
  }
}

bool line_compiled(int n) {
  switch (n) {
#define TM_CASE(k) case k:
    TM_LINE_SIZES(TM_CASE)
#undef TM_CASE
    return true;
    default: return false;
  }
}

// Bluestein tables of length N and sign s (chirp c_n = exp(s i pi n^2 / N), bhat = FFT_M(b) / M with
// b_m = conj(c_|m|), |m| < N), computed on the host in double and cached on the handle.
tm_refine::BluesteinTables& bluestein_tables(tm_refine* h, int N, int sign) {
  auto key = std::make_pair(N, sign);
  auto it = h->bluestein.find(key);
  if (it != h->bluestein.end()) return *it->second;
  auto bt = std::make_unique<tm_refine::BluesteinTables>();
  int M = 1;
  while (M < 2 * N - 1) M *= 2;
  TM_REQUIRE(M == 2048 || M == 4096, TM_ERR_UNSUPPORTED, "transform length " + std::to_string(N) + " not supported");
  bt->M = M;
  const long double pi = acosl(-1.0L);
  std::vector<cf> c(N);
  std::vector<std::complex<double>> b(M, 0.0), wt(M);
  for (long n = 0; n < N; ++n) {
    const long double a = sign * pi * (long double)((n * n) % (2L * N)) / (long double)N;
    c[n].x = (float)cosl(a);
    c[n].y = (float)sinl(a);
    const std::complex<double> cc((double)cosl(a), -(double)sinl(a));
    b[n] = cc;
    if (n > 0) b[M - n] = cc;
  }
  for (int k = 0; k < M; ++k) {
    const long double a = -2.0L * pi * (long double)k / (long double)M;
    wt[k] = std::complex<double>((double)cosl(a), (double)sinl(a));
  }
  std::vector<cf> bh(M);
  for (int k = 0; k < M; ++k) {
    std::complex<double> acc = 0.0;
    for (int m = 0; m < M; ++m)
      if (m < N || m > M - N) acc += b[m] * wt[(int)(((long long)m * k) % M)];
    bh[k].x = (float)(acc.real() / M);
    bh[k].y = (float)(acc.imag() / M);
  }
  bt->chirp.alloc(N);
  bt->bhat.alloc(M);
  TM_HIP(hipMemcpy(bt->chirp.p, c.data(), N * sizeof(cf), hipMemcpyHostToDevice));
  TM_HIP(hipMemcpy(bt->bhat.p, bh.data(), M * sizeof(cf), hipMemcpyHostToDevice));
  auto& ref = *bt;
  h->bluestein[key] = std::move(bt);
  return ref;
}

// Inverse column transforms (k_cols layout) of any length N.
void inverse_cols_any(tm_refine* h, int N, cf* data, int ncols, size_t bstride, int nb) {
  if (line_compiled(N)) { cols_line(h, N, data, ncols, bstride, nb, +1); return; }
  auto& bt = bluestein_tables(h, N, +1);
  const dim3 g((unsigned)((size_t)ncols * nb));
  if (bt.M == 2048)
    hipLaunchKernelGGL((tmk::k_bluestein_cols<2048, 256>), g, dim3(256), 0, h->stream, data, N, ncols, bstride, bt.chirp.p,
                       bt.bhat.p, h->twiddle(2048));
  else
    hipLaunchKernelGGL((tmk::k_bluestein_cols<4096, 256>), g, dim3(256), 0, h->stream, data, N, ncols, bstride, bt.chirp.p,
                       bt.bhat.p, h->twiddle(4096));
  check_launch("k_bluestein_cols");
}

// C2R rows of any even length N (in place allowed).
void c2r_rows_any(tm_refine* h, int N, cf* in, size_t ibs, int irs, float* out, size_t obs, int ors, int nrows, int nb) {
  if (template_size_compiled(N)) { rows_c2r(h, N, in, ibs, irs, out, obs, ors, nrows, nb); return; }
  auto& bt = bluestein_tables(h, N, +1);
  if (bt.M == 2048)
    hipLaunchKernelGGL((tmk::k_bluestein_c2r_rows<2048, 256>), dim3(nrows, nb), dim3(256), 0, h->stream, in, N, ibs, irs, out,
                       obs, ors, bt.chirp.p, bt.bhat.p, h->twiddle(2048));
  else
    hipLaunchKernelGGL((tmk::k_bluestein_c2r_rows<4096, 256>), dim3(nrows, nb), dim3(256), 0, h->stream, in, N, ibs, irs, out,
                       obs, ors, bt.chirp.p, bt.bhat.p, h->twiddle(4096));
  check_launch("k_bluestein_c2r_rows");
}

// TemplateScore's ChangePixelSize (refine_template.cpp:39, 49; image.cpp:8705-8792) for the items
// [first, first + count) of a batch: their a x a real projections (src, rows of a + 2 floats, item
// stride sstride floats) are resampled by factors[i] into t x t spectra (scaled like ForwardFFT)
// in their projection slots.
void resample_projections(tm_refine* h, const float* src, size_t sstride, int a, int first, int count,
                          const std::vector<float>& factors) {
  const int t = h->t, th = h->th;
  const size_t proj_bs = (size_t)t * th;
  const int K = h->K;
  constexpr int BIG = 1000, BH = BIG / 2 + 1;
  if (!h->rs_pad.p) {
    h->rs_pad.alloc(K);
    h->rs_big.alloc((size_t)K * BIG * (BIG + 2));
    h->rs_spec.alloc((size_t)K * BH * BIG);
    h->rs_clip.alloc((size_t)K * t * (t + 2));
  }
  int g0 = 0;
  while (g0 < count) {
    int pad, nw;
    change_pixel_size_dims(a, factors[g0], 0.001f, &pad, &nw);
    int m = 1;
    while (g0 + m < count) {
      int p2, n2;
      change_pixel_size_dims(a, factors[g0 + m], 0.001f, &p2, &n2);
      if (p2 != pad || n2 != nw) break;
      ++m;
    }
    TM_REQUIRE(pad == BIG, TM_ERR_UNSUPPORTED,
               "pixel-size change pads a " + std::to_string(a) + "^2 projection to " + std::to_string(pad) +
                   "; only the intermediate edge 1000 (the reference's tolerance for edges below 1000 / factor) is compiled");
    TM_REQUIRE(nw % 2 == 0 && nw <= 2048, TM_ERR_UNSUPPORTED, "pixel-size change to edge " + std::to_string(nw));
    const int nh = nw / 2 + 1;
    if (h->rs_new.n < (size_t)K * nh * nw) h->rs_new.alloc((size_t)K * nh * nw);
    const float* s0 = src + (size_t)g0 * sstride;
    hipLaunchKernelGGL((tmk::k_masked_mean2<256>), dim3(m), dim3(256), 0, h->stream, s0, a, sstride, 0.45f * a, h->rs_pad.p);
    check_launch("k_masked_mean2");
    hipLaunchKernelGGL(tmk::k_real_clip2, dim3(1024, m), dim3(256), 0, h->stream, s0, a, sstride, h->rs_big.p, BIG,
                       (size_t)BIG * (BIG + 2), h->rs_pad.p);
    check_launch("k_real_clip2");
    const float sbig = (float)(1. / (double)(float)((long long)BIG * BIG));
    rows_r2c(h, BIG, h->rs_big.p, (size_t)BIG * (BIG + 2), BIG + 2, 0, 0, BIG, 0, h->rs_spec.p, (size_t)BH * BIG, BH, BIG, m,
             sbig);
    cols_line(h, BIG, h->rs_spec.p, BH, (size_t)BH * BIG, m, -1);
    hipLaunchKernelGGL(tmk::k_fourier_clip2, dim3(512, m), dim3(256), 0, h->stream, h->rs_spec.p, BIG, (size_t)BH * BIG,
                       h->rs_new.p, nw, (size_t)nh * nw);
    check_launch("k_fourier_clip2");
    cf* dst = h->proj.p + (size_t)(first + g0) * proj_bs;
    if (nw == t) {
      // the new edge is the box's: the resampled spectrum is the projection (other_image->CopyFrom)
      hipLaunchKernelGGL(tmk::k_copy_scale_cf, dim3(64, m), dim3(256), 0, h->stream, h->rs_new.p, (size_t)nh * nw, dst,
                         proj_bs, (int)proj_bs, 1.0f);
      check_launch("k_copy_scale_cf");
    } else {
      inverse_cols_any(h, nw, h->rs_new.p, nh, (size_t)nh * nw, m);
      float* nr = reinterpret_cast<float*>(h->rs_new.p);
      c2r_rows_any(h, nw, h->rs_new.p, (size_t)nh * nw, nh, nr, (size_t)(nw + 2) * nw, nw + 2, nw, m);
      hipLaunchKernelGGL(tmk::k_real_clip2, dim3(64, m), dim3(256), 0, h->stream, nr, nw, (size_t)(nw + 2) * nw,
                         h->rs_clip.p, t, (size_t)t * (t + 2), h->rs_pad.p);
      check_launch("k_real_clip2");
      const float st = (float)(1. / (double)(float)((long long)t * t));
      rows_r2c(h, t, h->rs_clip.p, (size_t)t * (t + 2), t + 2, 0, 0, t, 0, dst, proj_bs, th, t, m, st);
      cols(h, t, dst, th, proj_bs, m, -1, 0, nullptr, 1.0f);
    }
    g0 += m;
  }
}

// One batch of TemplateScore calls: item b correlates the projection at reqs[b] with window slots[b].
// Items whose pixel-size factor is not within the tolerance of 1 go last, ordered by their resampled
// edge, and take the resampling path after the slice's C2R.
// The fused path split in two so that one batch's host work overlaps the previous batch on the
// device: launch_fused fills half `half` of the pinned and device request buffers, enqueues the
// copies, k_score_fused and the copy back, and records batch_done[half]; finish_fused waits for that
// event and returns the results. Requests keep their order (no pixel-size change to group by).
bool fused_applies(const tm_refine* h) { return h->fused && !h->timing && h->P == h->t && h->t == 64; }

void launch_fused(tm_refine* h, const std::vector<ScoreRequest>& reqs, const std::vector<int>& slot_in, int half) {
  const int n = (int)reqs.size(), t = h->t, K = h->K;
  TM_REQUIRE(n <= K, TM_ERR_INVALID, "batch larger than max_in_flight");
  float* hm = h->h_mats.p + (size_t)half * 9 * K;
  float* hd = h->h_dfs.p + (size_t)half * 2 * K;
  int* hs = h->h_slots.p + (size_t)half * K;
  for (int q = 0; q < n; ++q) {
    const ScoreRequest& r = reqs[q];
    euler_matrix(r.phi, r.theta, r.psi, &hm[9 * q]);
    hd[2 * q] = r.df1;
    hd[2 * q + 1] = r.df2;
    hs[q] = slot_in[q];
  }
  float* dm = h->mats.p + (size_t)half * 9 * K;
  float* dd = h->dfs.p + (size_t)half * 2 * K;
  int* ds = h->slots.p + (size_t)half * K;
  float* dr = h->results.p + (size_t)half * 3 * K;
  TM_HIP(hipMemcpyAsync(dm, hm, 9 * n * sizeof(float), hipMemcpyHostToDevice, h->stream));
  TM_HIP(hipMemcpyAsync(dd, hd, 2 * n * sizeof(float), hipMemcpyHostToDevice, h->stream));
  TM_HIP(hipMemcpyAsync(ds, hs, n * sizeof(int), hipMemcpyHostToDevice, h->stream));
  const float scale = (float)(1. / (double)(float)((long long)t * t));
  hipLaunchKernelGGL((rtk::k_score_fused<64>), dim3(n), dim3(256), 0, h->stream, h->tpl.p, dm, h->phaseT.p,
                     h->twiddle(t), h->twiddle(t / 2), h->win.p, ds, dd, h->ctf, h->curve_x.p, h->curve_y.p,
                     h->n_curve, scale, dr);
  check_launch("k_score_fused");
  TM_HIP(hipMemcpyAsync(h->h_results.p + (size_t)half * 3 * K, dr, 3 * n * sizeof(float), hipMemcpyDeviceToHost,
                        h->stream));
  if (!h->batch_done[half]) TM_HIP(hipEventCreateWithFlags(&h->batch_done[half], hipEventDisableTiming));
  TM_HIP(hipEventRecord(h->batch_done[half], h->stream));
}

void finish_fused(tm_refine* h, int half, int n, std::vector<ScoreResult>& out) {
  TM_HIP(hipEventSynchronize(h->batch_done[half]));
  const float* r = h->h_results.p + (size_t)half * 3 * h->K;
  out.resize(n);
  for (int q = 0; q < n; ++q) out[q] = ScoreResult{r[3 * q], r[3 * q + 1], r[3 * q + 2]};
}

void run_scores(tm_refine* h, const std::vector<ScoreRequest>& reqs_in, const std::vector<int>& slot_in,
                std::vector<ScoreResult>& out) {
  const int n = (int)reqs_in.size();
  const int t = h->t, th = h->th, P = h->P, Ph = h->Ph;
  TM_REQUIRE(n <= h->K, TM_ERR_INVALID, "batch larger than max_in_flight");
  std::vector<int> ord(n);
  for (int b = 0; b < n; ++b) ord[b] = b;
  auto resampled = [&](int b) { return !(fabsf(reqs_in[b].factor - 1.0f) < 0.001f); };
  auto new_edge = [&](int b) {
    int pad, nw;
    change_pixel_size_dims(P != t ? P : t, reqs_in[b].factor, 0.001f, &pad, &nw);
    return nw;
  };
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
    const bool rx = resampled(x), ry = resampled(y);
    if (rx != ry) return !rx;
    return rx && new_edge(x) < new_edge(y);
  });
  int n0 = 0;
  while (n0 < n && !resampled(ord[n0])) ++n0;
  std::vector<float> factors;
  for (int q = n0; q < n; ++q) factors.push_back(reqs_in[ord[q]].factor);
  for (int q = 0; q < n; ++q) {
    const ScoreRequest& r = reqs_in[ord[q]];
    euler_matrix(r.phi, r.theta, r.psi, &h->h_mats.p[9 * q]);
    h->h_dfs.p[2 * q] = r.df1;
    h->h_dfs.p[2 * q + 1] = r.df2;
    h->h_slots.p[q] = slot_in[ord[q]];
  }
  TM_HIP(hipMemcpyAsync(h->mats.p, h->h_mats.p, 9 * n * sizeof(float), hipMemcpyHostToDevice, h->stream));
  TM_HIP(hipMemcpyAsync(h->dfs.p, h->h_dfs.p, 2 * n * sizeof(float), hipMemcpyHostToDevice, h->stream));
  TM_HIP(hipMemcpyAsync(h->slots.p, h->h_slots.p, n * sizeof(int), hipMemcpyHostToDevice, h->stream));
  const size_t proj_bs = (size_t)t * th;
  const float scale = (float)(1. / (double)(float)((long long)t * t));
  using R = tm_refine;
  if (h->fused && P == t && t == 64 && n0 == n) {
    {
      Stage st(h, R::S_FUSED);
      hipLaunchKernelGGL((rtk::k_score_fused<64>), dim3(n), dim3(256), 0, h->stream, h->tpl.p, h->mats.p, h->phaseT.p,
                         h->twiddle(t), h->twiddle(t / 2), h->win.p, h->slots.p, h->dfs.p, h->ctf, h->curve_x.p,
                         h->curve_y.p, h->n_curve, scale, h->results.p);
      check_launch("k_score_fused");
    }
    TM_HIP(hipMemcpyAsync(h->h_results.p, h->results.p, 3 * n * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    TM_HIP(hipStreamSynchronize(h->stream));
    if (h->timing) {
      h->batch_items = n;
      h->resolve_timing();
    }
    out.resize(n);
    for (int q = 0; q < n; ++q)
      out[ord[q]] = ScoreResult{h->h_results.p[3 * q], h->h_results.p[3 * q + 1], h->h_results.p[3 * q + 2]};
    return;
  }
  if (P != t) {
    // ExtractSlice, SwapRealSpaceQuadrants, BackwardFFT (columns, then the t centre rows that
    // ClipInto keeps, or every row for a resampled item), ClipInto t^2 + ForwardFFT (rows, scaled
    // 1/N; columns below)
    const size_t slice_bs = (size_t)P * Ph;
    const int r0 = P / 2 - t / 2;
    {
      Stage st(h, R::S_EXTRACT);
      extract_cols<1>(h, P, h->tpl.p, h->mats.p, 0, n, h->phaseP.p, nullptr, h->slice.p);
    }
    {
      Stage st(h, R::S_SLICE_C2R);
      if (n0 > 0)
        rows_c2r(h, P, h->slice.p + (size_t)r0 * Ph, slice_bs, Ph,
                 reinterpret_cast<float*>(h->slice.p) + (size_t)r0 * (P + 2), 2 * slice_bs, P + 2, t, n0);
      if (n > n0)
        rows_c2r(h, P, h->slice.p + (size_t)n0 * slice_bs, slice_bs, Ph,
                 reinterpret_cast<float*>(h->slice.p + (size_t)n0 * slice_bs), 2 * slice_bs, P + 2, P, n - n0);
    }
    Stage st(h, R::S_PROJ_R2C);
    if (n0 > 0) {
      rows_r2c(h, t, reinterpret_cast<const float*>(h->slice.p), 2 * slice_bs, P + 2, r0, r0, t, 0, h->proj.p, proj_bs, th,
               t, n0, scale);
      cols(h, t, h->proj.p, th, proj_bs, n0, -1, 0, nullptr, 1.0f);
    }
    if (n > n0)
      resample_projections(h, reinterpret_cast<const float*>(h->slice.p + (size_t)n0 * slice_bs), 2 * slice_bs, P, n0,
                           n - n0, factors);
  } else {
    // ExtractSlice, SwapRealSpaceQuadrants, BackwardFFT, ForwardFFT (rows in place)
    {
      Stage st(h, R::S_EXTRACT);
      extract_cols<1>(h, t, h->tpl.p, h->mats.p, 0, n, h->phaseT.p, nullptr, h->proj.p);
    }
    {
      Stage st(h, R::S_SLICE_C2R);
      rows_c2r(h, t, h->proj.p, proj_bs, th, reinterpret_cast<float*>(h->proj.p), 2 * proj_bs, t + 2, t, n);
    }
    Stage st(h, R::S_PROJ_R2C);
    if (n0 > 0) {
      rows_r2c(h, t, reinterpret_cast<const float*>(h->proj.p), 2 * proj_bs, t + 2, 0, 0, t, 0, h->proj.p, proj_bs, th, t,
               n0, scale);
      cols(h, t, h->proj.p, th, proj_bs, n0, -1, 0, nullptr, 1.0f);
    }
    if (n > n0)
      resample_projections(h, reinterpret_cast<const float*>(h->proj.p + (size_t)n0 * proj_bs), 2 * proj_bs, t, n0, n - n0,
                           factors);
  }
  constexpr int NT = 256;
  {
    Stage st(h, R::S_PRODUCT);
    hipLaunchKernelGGL((rtk::k_score_product<NT>), dim3(h->score_parts, n), dim3(NT), 0, h->stream, h->proj.p, t, h->win.p,
                       h->slots.p, h->dfs.p, h->ctf, h->curve_x.p, h->curve_y.p, h->n_curve, h->partial.p);
    check_launch("k_score_product");
  }
  {
    Stage st(h, R::S_CCF_C2R);
    cols(h, t, h->proj.p, th, proj_bs, n, +1, 0, nullptr, 1.0f);
    rows_c2r(h, t, h->proj.p, proj_bs, th, reinterpret_cast<float*>(h->proj.p), 2 * proj_bs, t + 2, t, n);
  }
  {
    Stage st(h, R::S_PEAK);
    const int nq = h->peak_parts;
    hipLaunchKernelGGL((rtk::k_score_peak_rows<NT>), dim3(nq, n), dim3(NT), 0, h->stream,
                       reinterpret_cast<const float*>(h->proj.p), t, h->peak_v.p, h->peak_i.p);
    check_launch("k_score_peak_rows");
    hipLaunchKernelGGL(rtk::k_score_peak_final, dim3((n + 63) / 64), dim3(64), 0, h->stream, h->peak_v.p, h->peak_i.p, nq,
                       t, h->partial.p, h->score_parts, n, h->results.p);
    check_launch("k_score_peak_final");
  }
  TM_HIP(hipMemcpyAsync(h->h_results.p, h->results.p, 3 * n * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  TM_HIP(hipStreamSynchronize(h->stream));
  if (h->timing) {
    h->batch_items = n;
    h->resolve_timing();
  }
  out.resize(n);
  for (int q = 0; q < n; ++q)
    out[ord[q]] = ScoreResult{h->h_results.p[3 * q], h->h_results.p[3 * q + 1], h->h_results.p[3 * q + 2]};
}

struct FoundPeak {
  float x, y, value;
};

// The peak list (refine_template.cpp:561-601) on the device copy of the scaled MIP (consumed).
std::vector<FoundPeak> peak_list(tm_refine* h, long long cap) {
  const int mx = h->prm.image_x, my = h->prm.image_y;
  constexpr int NT = 256, NB = 1024;
  const float radius = h->prm.min_peak_radius;
  std::vector<FoundPeak> out;
  while ((long long)out.size() < cap) {
    hipLaunchKernelGGL((rtk::k_map_argmax<NT>), dim3(NB), dim3(NT), 0, h->stream, h->smip.p, mx, my, 50, h->arg_v.p,
                       h->arg_i.p);
    check_launch("k_map_argmax");
    hipLaunchKernelGGL((rtk::k_argmax_final<NT>), dim3(1), dim3(NT), 0, h->stream, h->arg_v.p, h->arg_i.p, NB,
                       h->arg_v.p + NB, h->arg_i.p + NB);
    check_launch("k_argmax_final");
    TM_HIP(hipMemcpyAsync(h->h_arg_v.p, h->arg_v.p + NB, sizeof(float), hipMemcpyDeviceToHost, h->stream));
    TM_HIP(hipMemcpyAsync(h->h_arg_i.p, h->arg_i.p + NB, sizeof(long long), hipMemcpyDeviceToHost, h->stream));
    TM_HIP(hipStreamSynchronize(h->stream));
    const float v = h->h_arg_v.p[0];
    const long long e = h->h_arg_i.p[0];
    FoundPeak pk{0.0f, 0.0f, v};
    if (e != LLONG_MAX) {
      pk.x = (float)((int)(e % mx) - mx / 2);
      pk.y = (float)((int)(e / mx) - my / 2);
    }
    if (pk.value < h->prm.wanted_threshold) break;
    out.push_back(pk);
    const float cx = pk.x + (float)(mx / 2), cy = pk.y + (float)(my / 2);
    if (radius >= 0.0f) {
      const double rr = std::sqrt((double)radius) + 2.0;
      const int x0 = std::max(0, (int)std::floor(cx - rr)), x1 = std::min(mx - 1, (int)std::ceil(cx + rr));
      const int y0 = std::max(0, (int)std::floor(cy - rr)), y1 = std::min(my - 1, (int)std::ceil(cy + rr));
      const int w = x1 - x0 + 1, hg = y1 - y0 + 1;
      if (w > 0 && hg > 0) {
        const long long cnt = (long long)w * hg;
        hipLaunchKernelGGL(rtk::k_map_mask, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, h->stream, h->smip.p, mx,
                           x0, y0, w, hg, cx, cy, radius);
        check_launch("k_map_mask");
      }
    }
  }
  return out;
}

void upload_scaled_mip(tm_refine* h, const float* scaled_mip) {
  const size_t n = (size_t)h->prm.image_x * h->prm.image_y;
  TM_HIP(hipMemcpyAsync(h->smip.p, scaled_mip, n * sizeof(float), hipMemcpyHostToDevice, h->stream));
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

const char* tm_refine_last_error(void) { return g_rerr.c_str(); }

int tm_refine_create(int device_ordinal, tm_refine** out) {
  return rguard([&] {
    TM_REQUIRE(out, TM_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int n = 0;
    TM_HIP(hipGetDeviceCount(&n));
    TM_REQUIRE(device_ordinal >= 0 && device_ordinal < n, TM_ERR_INVALID,
               "device ordinal " + std::to_string(device_ordinal) + " out of range (" + std::to_string(n) + " devices)");
    TM_HIP(hipSetDevice(device_ordinal));
    auto* h = new tm_refine();
    h->device = device_ordinal;
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete h;
      throw TmError{TM_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)};
    }
    *out = h;
  });
}

int tm_refine_destroy(tm_refine* h) {
  return rguard([&] {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    delete h;
  });
}

int tm_refine_init(tm_refine* h, const tm_refine_params* p, const float* whitened_image, const float* template_spectrum,
                   const float* curve_x, const float* curve_y, int n_curve) {
  return rguard([&] {
    TM_REQUIRE(h && p && whitened_image && template_spectrum && curve_x && curve_y, TM_ERR_INVALID, "NULL argument");
    TM_HIP(hipSetDevice(h->device));
    h->initialised = false;
    const int mx = p->image_x, my = p->image_y, t = p->template_size, P = p->padded_size;
    TM_REQUIRE(mx > 0 && my > 0 && mx % 2 == 0, TM_ERR_INVALID, "image_x must be positive and even, image_y positive");
    TM_REQUIRE(template_size_compiled(t) && template_size_compiled(P) && P >= t, TM_ERR_UNSUPPORTED,
               "template size " + std::to_string(t) + " / padded size " + std::to_string(P) +
                   " not supported (powers of two 32..1024, padded >= template)");
    TM_REQUIRE(n_curve >= 2, TM_ERR_INVALID, "the whitening curve needs at least two points");
    TM_REQUIRE(p->pixel_size > 0.0f, TM_ERR_INVALID, "pixel_size must be positive");
    TM_REQUIRE(p->max_in_flight >= 0, TM_ERR_INVALID, "max_in_flight must be >= 0");
    h->prm = *p;
    h->t = t;
    h->th = t / 2 + 1;
    h->P = P;
    h->Ph = P / 2 + 1;
    h->K = p->shared_local_maps ? 1 : (p->max_in_flight > 0 ? p->max_in_flight : 256);
    h->n_curve = n_curve;
    // CTF::Init (ctf.cpp:93-150) in pixel units
    const double PI = 3.14159265359;
    const float wl =
        12.2639f / sqrtf(1000.0f * p->voltage_kV + 0.97845e-6 * powf(1000.0f * p->voltage_kV, 2)) / p->pixel_size;
    h->ctf.wavelength = wl;
    h->ctf.squared_wavelength = powf(wl, 2);
    h->ctf.spherical_aberration = p->spherical_aberration_mm * 10000000.0 / p->pixel_size;
    h->ctf.phase_shift = (float)(p->phase_shift_deg * 3.14159265359f / 180.);
    h->ctf.amplitude_term = (fabs(p->amplitude_contrast - 1.0) < 1e-3)
                                ? (float)(PI / 2.0f)
                                : atanf(p->amplitude_contrast / sqrtf(1.0 - powf(p->amplitude_contrast, 2)));
    h->ctf.astigmatism_azimuth = (float)(p->defocus_angle * 3.14159265359f / 180.);
    const size_t npx = (size_t)mx * my;
    h->img.alloc(npx);
    TM_HIP(hipMemcpy(h->img.p, whitened_image, npx * sizeof(float), hipMemcpyHostToDevice));
    h->smip.alloc(npx);
    {
      // the template spectrum in the bricked layout the extraction reads (tpl_index<1>)
      DevBuf<cf> linear;
      linear.alloc((size_t)P * P * h->Ph);
      TM_HIP(hipMemcpy(linear.p, template_spectrum, linear.n * sizeof(cf), hipMemcpyHostToDevice));
      h->tpl.alloc((size_t)(P / 2) * (P / 2) * ((P / 2 + 4) / 4) * 16);
      hipLaunchKernelGGL(tmk::k_tpl_to_bricks, dim3(4096), dim3(256), 0, h->stream, linear.p, P, h->tpl.p);
      check_launch("k_tpl_to_bricks");
      TM_HIP(hipStreamSynchronize(h->stream));
    }
    h->curve_x.alloc(n_curve);
    h->curve_y.alloc(n_curve);
    TM_HIP(hipMemcpy(h->curve_x.p, curve_x, n_curve * sizeof(float), hipMemcpyHostToDevice));
    TM_HIP(hipMemcpy(h->curve_y.p, curve_y, n_curve * sizeof(float), hipMemcpyHostToDevice));
    auto pp = swap_phase(P);
    h->phaseP.alloc(pp.size());
    TM_HIP(hipMemcpy(h->phaseP.p, pp.data(), pp.size() * sizeof(cf), hipMemcpyHostToDevice));
    auto pt = swap_phase(t);
    h->phaseT.alloc(pt.size());
    TM_HIP(hipMemcpy(h->phaseT.p, pt.data(), pt.size() * sizeof(cf), hipMemcpyHostToDevice));
    const int K = h->K;
    const size_t per = (size_t)t * h->th;
    h->score_parts = (int)std::max<size_t>(1, std::min<size_t>(32, per / 2048));
    h->mats.alloc(2 * 9 * (size_t)K);  // two halves: the fused path overlaps batches
    h->dfs.alloc(2 * 2 * (size_t)K);
    h->results.alloc(2 * 3 * (size_t)K);
    h->slots.alloc(2 * (size_t)K);
    h->shifts.alloc(K);
    h->partial.alloc((size_t)K * h->score_parts);
    h->peak_parts = std::max(1, std::min(t, t * t / 4096));
    h->peak_v.alloc((size_t)K * h->peak_parts);
    h->peak_i.alloc((size_t)K * h->peak_parts);
    if (P != t) h->slice.alloc((size_t)K * P * h->Ph);
    h->proj.alloc((size_t)K * per);
    h->win.alloc((size_t)K * per);
    h->win_stage.alloc((size_t)K * per);
    h->win_real.alloc((size_t)K * t * (t + 2));
    h->mask_w.release();
    h->mask_band = 0;
    if (p->mask_radius > 0.0f) {
      // refine_template.cpp:520-521 (the radius is at most half the template box less one pixel),
      // :461 (20 A falloff) and :716 (radius and falloff in pixels)
      float mask_radius = p->mask_radius;
      const float temp_float = (float(t) / 2.0f - 1.0f) * p->pixel_size;
      if (mask_radius > temp_float) mask_radius = temp_float;
      const float mask_falloff = 20.0;
      int nb = 0;
      auto w = mask_weights(t, mask_radius / p->pixel_size, mask_falloff / p->pixel_size, &nb);
      TM_REQUIRE(nb > 0, TM_ERR_INVALID, "the particle mask's edge band holds no pixel (falloff below one pixel)");
      h->mask_w.alloc(w.size());
      TM_HIP(hipMemcpy(h->mask_w.p, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
      h->mask_band = nb;
    }
    h->arg_v.alloc(1024 + 1);
    h->arg_i.alloc(1024 + 1);
    h->h_mats.alloc(2 * 9 * (size_t)K);
    h->h_dfs.alloc(2 * 2 * (size_t)K);
    h->h_results.alloc(2 * 3 * (size_t)K);
    h->h_slots.alloc(2 * (size_t)K);
    h->h_shifts.alloc(K);
    h->h_arg_v.alloc(1);
    h->h_arg_i.alloc(1);
    for (int n : {t, t / 2, P, P / 2}) h->twiddle(n);
    TM_HIP(hipDeviceSynchronize());
    h->initialised = true;
  });
}

int tm_refine_set_image(tm_refine* h, const float* whitened_image, const float* curve_x, const float* curve_y,
                        int n_curve) {
  return rguard([&] {
    require_ready(h);
    TM_REQUIRE(whitened_image && curve_x && curve_y, TM_ERR_INVALID, "NULL argument");
    TM_REQUIRE(n_curve == h->n_curve, TM_ERR_INVALID, "the whitening curve length differs from the handle's");
    const size_t npx = (size_t)h->prm.image_x * h->prm.image_y;
    TM_HIP(hipMemcpyAsync(h->img.p, whitened_image, npx * sizeof(float), hipMemcpyHostToDevice, h->stream));
    TM_HIP(hipMemcpyAsync(h->curve_x.p, curve_x, n_curve * sizeof(float), hipMemcpyHostToDevice, h->stream));
    TM_HIP(hipMemcpyAsync(h->curve_y.p, curve_y, n_curve * sizeof(float), hipMemcpyHostToDevice, h->stream));
    TM_HIP(hipStreamSynchronize(h->stream));
  });
}

int tm_refine_set_micrograph(tm_refine* h, const float* micrograph) {
  return rguard([&] {
    require_ready(h);
    TM_REQUIRE(micrograph, TM_ERR_INVALID, "micrograph is NULL");
    const int mx = h->prm.image_x, my = h->prm.image_y, nh = mx / 2 + 1, n = h->n_curve;
    TM_REQUIRE(line_compiled(mx) && line_compiled(my), TM_ERR_UNSUPPORTED,
               "no compiled transform for a " + std::to_string(mx) + " x " + std::to_string(my) + " micrograph");
    TM_REQUIRE(n == int((mx / 2.0 + 1.0) * sqrtf(2.0) + 1.0) && n >= 2, TM_ERR_INVALID,
               "the handle's whitening curve length does not match the micrograph");
    const long long npx = (long long)mx * my, nspec = (long long)nh * my;
    hipStream_t st = h->stream;
    if (h->pr_src.n != (size_t)npx) h->pr_src.alloc(npx);
    if (h->pr_pad.n != (size_t)(mx + 2) * my) h->pr_pad.alloc((size_t)(mx + 2) * my);
    if (h->pr_spec.n != (size_t)nspec) h->pr_spec.alloc(nspec);
    if (!h->pr_part.p) h->pr_part.alloc(2 * 1024);
    if (!h->pr_st.p) h->pr_st.alloc(3);
    if (h->pr_acc.n != (size_t)2 * n) h->pr_acc.alloc(2 * (size_t)n);
    TM_HIP(hipMemcpyAsync(h->pr_src.p, micrograph, npx * sizeof(float), hipMemcpyHostToDevice, st));
    // ReplaceOutliersWithMean(5)
    const long long chunk = std::max<long long>(256, (npx + 1023) / 1024);
    const int np = (int)((npx + chunk - 1) / chunk);
    hipLaunchKernelGGL(tmp::k_moments<256>, dim3(np), dim3(256), 0, st, h->pr_src.p, npx, chunk, h->pr_part.p);
    check_launch("k_moments");
    hipLaunchKernelGGL(tmp::k_outlier_bounds, dim3(1), dim3(1), 0, st, h->pr_part.p, np, npx, h->pr_st.p);
    check_launch("k_outlier_bounds");
    const unsigned g = (unsigned)std::min<long long>(65535, ((long long)(mx + 2) * my + 255) / 256);
    hipLaunchKernelGGL(tmp::k_outlier_pad, dim3(g), dim3(256), 0, st, h->pr_src.p, mx, my, h->pr_st.p, h->pr_pad.p);
    check_launch("k_outlier_pad");
    // ForwardFFT (scaled 1/N), ZeroCentralPixel
    rows_r2c(h, mx, h->pr_pad.p, (size_t)(mx + 2) * my, mx + 2, 0, 0, mx, 0, h->pr_spec.p, nspec, nh, my, 1, 1.0f);
    cols_line(h, my, h->pr_spec.p, nh, nspec, 1, -1);
    const float inv_n = 1. / (float)((long)mx * my);
    const unsigned gs = (unsigned)std::min<long long>(65535, (nspec + 255) / 256);
    hipLaunchKernelGGL(tmp::k_scale_zero_dc, dim3(gs), dim3(256), 0, st, h->pr_spec.p, nspec, inv_n);
    check_launch("k_scale_zero_dc");
    // the whitening curve on the curve's x axis (Curve::SetupXAxis(0, 0.5 sqrt 2, n))
    std::vector<float> x(n);
    const float lo = 0.0f, hi = 0.5 * sqrtf(2.0);
    for (int c = 0; c < n; ++c) x[c] = c * (hi - lo) / float(n - 1) + lo;
    TM_HIP(hipMemcpyAsync(h->curve_x.p, x.data(), n * sizeof(float), hipMemcpyHostToDevice, st));
    TM_HIP(hipMemsetAsync(h->pr_acc.p, 0, 2 * (size_t)n * sizeof(double), st));
    tmp::power_curve(st, h->pr_spec.p, mx, my, h->curve_x.p, n, h->pr_acc.p, h->pr_acc.p + n, gs);
    check_launch("k_power_curve");
    hipLaunchKernelGGL(tmp::k_curve_finish, dim3(1), dim3(256), 0, st, h->pr_acc.p, h->pr_acc.p + n, n, h->curve_y.p);
    check_launch("k_curve_finish");
    // ApplyCurveFilter, ZeroCentralPixel, DivideByConstant(sqrt(ReturnSumOfSquares()))
    hipLaunchKernelGGL(tmp::k_apply_curve, dim3(gs), dim3(256), 0, st, h->pr_spec.p, mx, my, h->curve_x.p,
                       h->curve_y.p, n);
    check_launch("k_apply_curve");
    const long long schunk = std::max<long long>(256, (nspec + 1023) / 1024);
    const int nsp = (int)((nspec + schunk - 1) / schunk);
    hipLaunchKernelGGL(tmp::k_spectrum_ssq<256>, dim3(nsp), dim3(256), 0, st, h->pr_spec.p, mx, my, schunk,
                       h->pr_part.p);
    check_launch("k_spectrum_ssq");
    hipLaunchKernelGGL(tmp::k_normalise_spectrum, dim3(gs), dim3(256), 0, st, h->pr_spec.p, nspec, h->pr_part.p, nsp,
                       0);
    check_launch("k_normalise_spectrum");
    // BackwardFFT (unscaled) into the handle's whitened image
    cols_line(h, my, h->pr_spec.p, nh, nspec, 1, +1);
    rows_c2r_line(h, mx, h->pr_spec.p, nspec, nh, h->pr_pad.p, (size_t)(mx + 2) * my, mx + 2, my, 1);
    hipLaunchKernelGGL(tmp::k_unpad, dim3((unsigned)std::min<long long>(65535, (npx + 255) / 256)), dim3(256), 0, st,
                       h->pr_pad.p, mx, my, h->img.p);
    check_launch("k_unpad");
    TM_HIP(hipStreamSynchronize(st));
  });
}

int tm_refine_get_image(tm_refine* h, float* whitened, float* curve_x, float* curve_y) {
  return rguard([&] {
    require_ready(h);
    TM_HIP(hipStreamSynchronize(h->stream));
    const size_t npx = (size_t)h->prm.image_x * h->prm.image_y;
    if (whitened) TM_HIP(hipMemcpy(whitened, h->img.p, npx * sizeof(float), hipMemcpyDeviceToHost));
    if (curve_x) TM_HIP(hipMemcpy(curve_x, h->curve_x.p, h->n_curve * sizeof(float), hipMemcpyDeviceToHost));
    if (curve_y) TM_HIP(hipMemcpy(curve_y, h->curve_y.p, h->n_curve * sizeof(float), hipMemcpyDeviceToHost));
  });
}

int tm_refine_peak_list(tm_refine* h, const float* scaled_mip, float* out, int capacity, int* n_peaks) {
  return rguard([&] {
    require_ready(h);
    TM_REQUIRE(scaled_mip && n_peaks && (out || capacity == 0), TM_ERR_INVALID, "NULL argument");
    upload_scaled_mip(h, scaled_mip);
    auto pk = peak_list(h, (long long)h->prm.image_x * h->prm.image_y / 100);
    for (int q = 0; q < (int)pk.size() && q < capacity; ++q) {
      out[3 * q] = pk[q].x;
      out[3 * q + 1] = pk[q].y;
      out[3 * q + 2] = pk[q].value;
    }
    *n_peaks = (int)pk.size();
  });
}

int tm_refine_score(tm_refine* h, int peak_x, int peak_y, float current_defocus, float step, float phi, float theta,
                    float psi, float pixel_size_factor, float* out) {
  return rguard([&] {
    require_ready(h);
    TM_REQUIRE(out, TM_ERR_INVALID, "out is NULL");
    make_windows(h, {peak_x}, {peak_y}, {0});
    const tm_refine_params& p = h->prm;
    std::vector<ScoreResult> res;
    run_scores(h, {ScoreRequest{phi, theta, psi, (p.defocus1 + current_defocus + step) / p.pixel_size,
                                (p.defocus2 + current_defocus + step) / p.pixel_size, pixel_size_factor}},
               {0}, res);
    out[0] = res[0].x;
    out[1] = res[0].y;
    out[2] = res[0].value;
  });
}

int tm_refine_score_batch(tm_refine* h, int n, const int* peak_xy, const float* angles, float current_defocus,
                          float* out) {
  return rguard([&] {
    require_ready(h);
    TM_REQUIRE(n >= 0 && (n == 0 || (peak_xy && angles && out)), TM_ERR_INVALID, "NULL argument");
    // windows in order of first appearance, each with its requests in input order
    std::vector<std::pair<int, int>> wins;
    std::vector<std::vector<int>> reqs_of;
    std::map<std::pair<int, int>, int> index;
    for (int i = 0; i < n; ++i) {
      const auto key = std::make_pair(peak_xy[2 * i], peak_xy[2 * i + 1]);
      auto it = index.find(key);
      if (it == index.end()) {
        it = index.emplace(key, (int)wins.size()).first;
        wins.push_back(key);
        reqs_of.emplace_back();
      }
      reqs_of[it->second].push_back(i);
    }
    const tm_refine_params& p = h->prm;
    const float df1 = (p.defocus1 + current_defocus) / p.pixel_size, df2 = (p.defocus2 + current_defocus) / p.pixel_size;
    std::vector<ScoreRequest> reqs;
    std::vector<int> slots, which;
    std::vector<ScoreResult> res;
    auto store = [&](const std::vector<int>& who) {
      for (size_t q = 0; q < who.size(); ++q) {
        out[3 * who[q]] = res[q].x;
        out[3 * who[q] + 1] = res[q].y;
        out[3 * who[q] + 2] = res[q].value;
      }
    };
    // fused path: batch k is prepared and enqueued while batch k - 1 runs, then k - 1 is collected
    const bool pipelined = fused_applies(h);
    int half = 0, pending_half = -1;
    std::vector<int> pending_which;
    auto drain = [&] {
      if (pending_half < 0) return;
      finish_fused(h, pending_half, (int)pending_which.size(), res);
      store(pending_which);
      pending_half = -1;
    };
    auto flush = [&] {
      if (reqs.empty()) return;
      if (pipelined) {
        launch_fused(h, reqs, slots, half);
        drain();
        pending_half = half;
        pending_which.swap(which);
        half ^= 1;
      } else {
        run_scores(h, reqs, slots, res);
        store(which);
      }
      reqs.clear();
      slots.clear();
      which.clear();
    };
    // on an error with a batch in flight, let its copies finish before the pinned buffers are reused
    struct Settle {
      tm_refine* h;
      const int* pending;
      ~Settle() {
        if (*pending >= 0) (void)hipStreamSynchronize(h->stream);
      }
    } settle{h, &pending_half};
    for (size_t w0 = 0; w0 < wins.size(); w0 += h->K) {
      const size_t w1 = std::min(wins.size(), w0 + (size_t)h->K);
      std::vector<int> xs, ys, ss;
      for (size_t w = w0; w < w1; ++w) {
        xs.push_back(wins[w].first);
        ys.push_back(wins[w].second);
        ss.push_back((int)(w - w0));
      }
      drain();  // make_windows reuses the pinned slot buffer and the window slots
      make_windows(h, xs, ys, ss);
      for (size_t w = w0; w < w1; ++w)
        for (int i : reqs_of[w]) {
          reqs.push_back(ScoreRequest{angles[3 * i], angles[3 * i + 1], angles[3 * i + 2], df1, df2, 1.0f});
          slots.push_back((int)(w - w0));
          which.push_back(i);
          if ((int)reqs.size() == h->K) flush();
        }
      flush();  // the next group of windows reuses the slots
    }
    drain();
  });
}

int tm_refine_set_fused(tm_refine* h, int enabled) {
  return rguard([&] {
    TM_REQUIRE(h, TM_ERR_INVALID, "handle is NULL");
    h->fused = enabled != 0;
  });
}

int tm_refine_set_timing(tm_refine* h, int enabled) {
  return rguard([&] {
    TM_REQUIRE(h, TM_ERR_INVALID, "handle is NULL");
    h->timing = enabled != 0;
  });
}

int tm_refine_get_timing(tm_refine* h, const char** names, double* ms, long long* launches, long long* items, int cap) {
  if (!h) {
    g_rerr = "handle is NULL";
    return TM_ERR_INVALID;
  }
  int n = 0;
  for (int k = 0; k < tm_refine::S_COUNT && n < cap; ++k, ++n) {
    if (names) names[n] = kStageNames[k];
    if (ms) ms[n] = h->stage_ms[k];
    if (launches) launches[n] = h->stage_launches[k];
    if (items) items[n] = h->stage_items[k];
  }
  return n;
}

int tm_refine_run(tm_refine* h, const float* const* maps, float* const* out_maps, tm_refined_peak* peaks, int capacity,
                  int* n_peaks) {
  return rguard([&] {
    require_ready(h);
    TM_REQUIRE(maps && n_peaks && (peaks || capacity == 0), TM_ERR_INVALID, "NULL argument");
    for (int k = 0; k < 7; ++k) TM_REQUIRE(maps[k], TM_ERR_INVALID, "NULL map");
    const tm_refine_params& p = h->prm;
    const int mx = p.image_x, my = p.image_y;
    const size_t npx = (size_t)mx * my;
    RunConst K;
    K.p = p;
    K.mip = maps[0];
    K.scaled_mip = maps[1];
    K.defocus_search_range = p.defocus_search_range;
    K.defocus_refine_step = 2.0f * p.defocus_search_step;
    if (K.defocus_refine_step <= 0.0) {
      K.defocus_search_range = 0.0f;
      K.defocus_refine_step = 100.0f;
    }
    K.defocus_step = std::max(K.defocus_refine_step, 100.0f);
    K.pixel_size_search_range = p.pixel_size_search_range;
    K.pixel_size_refine_step = 2.0f * p.pixel_size_search_step;
    if (K.pixel_size_refine_step <= 0.0) {
      K.pixel_size_search_range = 0.0f;
      K.pixel_size_refine_step = 100.0f;
    }
    K.sqrt_t = sqrtf(float(h->t) * float(h->t));
    K.rs = mx + 2;
    for (int k = 0; k < tm_refine::S_COUNT; ++k) h->stage_ms[k] = 0.0, h->stage_launches[k] = 0, h->stage_items[k] = 0;
    upload_scaled_mip(h, maps[1]);
    const auto found = peak_list(h, (long long)npx / 100);
    const int npk = (int)found.size();
    std::vector<PeakState> states(npk);
    std::vector<std::unique_ptr<LocalMaps>> own;
    LocalMaps shared;
    auto init_maps = [&](LocalMaps& L) {
      L.in[M_PSI] = maps[2];
      L.in[M_THETA] = maps[3];
      L.in[M_PHI] = maps[4];
      L.in[M_DEFOCUS] = maps[5];
      L.in[M_PIXEL_SIZE] = maps[6];
      L.mx = mx;
      L.my = my;
    };
    init_maps(shared);
    // up to K peaks in flight; slot s holds the window of the peak in flight there
    struct Flight {
      int peak = -1;
      std::unique_ptr<DescentTask> task;
    };
    std::vector<Flight> slot(h->K);
    int next = 0, active = 0;
    std::vector<ScoreRequest> reqs;
    std::vector<int> req_slot;
    std::vector<ScoreResult> res;
    while (true) {
      // admit peaks into free slots, windows first, then run each to its first TemplateScore call
      std::vector<int> xs, ys, ws;
      std::vector<int> admitted;
      for (int s = 0; s < h->K && next < npk; ++s) {
        if (slot[s].peak >= 0) continue;
        PeakState& st = states[next];
        st.x = found[next].x;
        st.y = found[next].y;
        st.value = found[next].value;
        if (p.shared_local_maps) {
          st.maps = &shared;
        } else {
          own.emplace_back(new LocalMaps());
          init_maps(*own.back());
          st.maps = own.back().get();
        }
        slot[s].peak = next;
        xs.push_back((int)st.x);
        ys.push_back((int)st.y);
        ws.push_back(s);
        admitted.push_back(s);
        ++next;
        ++active;
      }
      make_windows(h, xs, ys, ws);
      for (int s : admitted) {
        slot[s].task.reset(new DescentTask(refine_peak(K, states[slot[s].peak])));
        slot[s].task->h.resume();
      }
      // retire finished peaks (their records are complete)
      bool retired = false;
      for (int s = 0; s < h->K; ++s) {
        if (slot[s].peak < 0 || !slot[s].task->h.done()) continue;
        if (slot[s].task->h.promise().error) std::rethrow_exception(slot[s].task->h.promise().error);
        if (!p.shared_local_maps) own[slot[s].peak].reset();
        slot[s].task.reset();
        slot[s].peak = -1;
        --active;
        retired = true;
      }
      if (retired && next < npk) continue;
      if (active == 0) break;
      reqs.clear();
      req_slot.clear();
      for (int s = 0; s < h->K; ++s) {
        if (slot[s].peak < 0) continue;
        reqs.push_back(slot[s].task->h.promise().req);
        req_slot.push_back(s);
      }
      run_scores(h, reqs, req_slot, res);
      for (size_t q = 0; q < reqs.size(); ++q) {
        auto& c = slot[req_slot[q]].task->h;
        c.promise().res = res[q];
        c.resume();
      }
    }
    // output maps (refine_template.cpp:641-648, 1092-1132), peaks applied in peak-list order
    if (out_maps) {
      for (int k = 0; k < 7; ++k) {
        if (!out_maps[k]) continue;
        if (k < 2) std::fill(out_maps[k], out_maps[k] + npx, 0.0f);
        else std::memcpy(out_maps[k], maps[k], npx * sizeof(float));
      }
      for (int q = 0; q < npk; ++q) {
        const tm_refined_peak& r = states[q].rec;
        if (!(r.offset * p.pixel_size <= p.xy_change_threshold || !p.exclude_above_xy_threshold)) continue;
        const long long row = r.best_address / K.rs, col = r.best_address % K.rs;
        if (col >= mx) continue;
        const size_t i = (size_t)row * mx + col;
        const float v[7] = {r.mip, r.peak_height, r.r_psi, r.r_theta, r.r_phi, r.r_defocus, r.r_pixel_size};
        for (int k = 0; k < 7; ++k)
          if (out_maps[k]) out_maps[k][i] = v[k];
      }
    }
    for (int q = 0; q < npk && q < capacity; ++q) peaks[q] = states[q].rec;
    *n_peaks = npk;
  });
}

}  // extern "C"
