// Hot-path kernels of the template-matching inner loop (gfx950).
//
// Per batch of B orientations (one tm_run_inner_loop call runs ceil(n_orient / B) batches):
//   k_extract_cols Fourier slice of the P^3 template at R(phi,theta,psi), trilinear
//                  (src/core/image.cpp:2450-2607, 2899-3020), times the quadrant-swap phase and,
//                  when P == T, the projection filter (match_template.cpp:1007-1023), then the
//                  inverse column FFT of the slice, in LDS
//   k_cols / k_rows_c2r / k_rows_r2c
//                  the small 2-D transforms of the projection (C2R P^2, clip T^2, R2C T^2,
//                  x filter, C2R T^2; match_template.cpp:1009-1030)
//   k_norm_stats, k_norm_scale  edge mean, variance to the padded size, 1/sqrt(var) (match_template.cpp:1060-1068)
//   k_rows_r2c     forward row FFT of the T non-zero rows of the W x H padded reference
//                  (ClipIntoLargerRealSpace2D + first half of the R2C, image.cpp:8398-8464)
//   k_mid          per column tile: forward column FFT, DC zero, conj(ref) * input, inverse
//                  column FFT (match_template.cpp:1071-1088); writes the inter-pass plane Z
//   k_last         per row: C2R row FFT of Z for every orientation of the batch, then the MIP /
//                  argmax / FP64 sum / FP64 sum-of-squares / histogram update with the state of the
//                  row held in registers across the batch (match_template.cpp:1100-1149)
#pragma once

#include <cfloat>
#include <cstdint>

#include "tm_fft.hpp"

namespace tmk {

__device__ __forceinline__ cf cmul(cf a, cf b, cf c) { return cmul(a, cmul(b, c)); }

// LDS index padding: one slot per 16 keeps stride-16 Stockham stores off a single bank group.
__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }
__host__ __device__ constexpr int lds_len(int n) { return n + n / 16; }

// In-place Stockham FFT of one line in LDS (padded index), executed by NT threads (tid < NT)
// of a block in which every thread reaches the barriers. Each stage loads all butterfly inputs
// into registers, synchronises, then stores.
template <int N, int SIGN, int NT, int LR = 4, int S = 0, typename TW = const cf*>
__device__ __forceinline__ void fft_inplace(cf* buf, TW tw, int tid) {
  static_assert(fft_size_ok(N), "FFT length must be 2^a 3^b 5^c");
  if constexpr (S < plan_stages(N, LR)) {
    constexpr int R = plan_radix(N, S, LR);
    constexpr int NS = plan_ns(N, S, LR);
    constexpr int M = N / R;
    constexpr int CNT = (M + NT - 1) / NT;
    cf v[CNT][R];
    // lpad(i + 16 m) = lpad(i) + 17 m for i >= 0: with M (reads) or NS (writes) a multiple of 16 the
    // R addresses of a butterfly are one padded base plus compile-time offsets
#pragma unroll
    for (int c = 0; c < CNT; ++c) {
      const int j = tid + c * NT;
      if (tid < NT && j < M) {
        if constexpr (M % 16 == 0) {
          const cf* rd = buf + lpad(j);
#pragma unroll
          for (int q = 0; q < R; ++q) v[c][q] = rd[q * (M + M / 16)];
        } else {
#pragma unroll
          for (int q = 0; q < R; ++q) v[c][q] = buf[lpad(j + q * M)];
        }
      }
    }
    lds_barrier();
#pragma unroll
    for (int c = 0; c < CNT; ++c) {
      const int j = tid + c * NT;
      if (tid < NT && j < M) {
        const int k = j % NS;
        if constexpr (NS > 1) {
          cf w[R];
          stage_twiddles<SIGN, N, NS, R>(tw, k, w);
#pragma unroll
          for (int q = 1; q < R; ++q) v[c][q] = cmul(v[c][q], w[q]);
        }
        dftR<R, SIGN>(v[c]);
        const int base = (j / NS) * NS * R + k;
        if constexpr (NS % 16 == 0) {
          cf* wr = buf + lpad(base);
#pragma unroll
          for (int q = 0; q < R; ++q) wr[q * (NS + NS / 16)] = v[c][q];
        } else if constexpr (16 % NS == 0 && R > 16 / NS) {
          // q NS = (q mod PB) NS + 16 (q / PB) with PB = 16 / NS: PB padded bases, offsets 17 (q / PB)
          constexpr int PB = 16 / NS;
          cf* wr[PB];
#pragma unroll
          for (int r = 0; r < PB; ++r) wr[r] = buf + lpad(base + r * NS);
#pragma unroll
          for (int q = 0; q < R; ++q) wr[q % PB][17 * (q / PB)] = v[c][q];
        } else {
#pragma unroll
          for (int q = 0; q < R; ++q) buf[lpad(base + q * NS)] = v[c][q];
        }
      }
    }
    lds_barrier();
    fft_inplace<N, SIGN, NT, LR, S + 1, TW>(buf, tw, tid);
  }
}

// Work item of block bid out of n. Workgroups are dispatched round-robin over the 8 XCDs, so XCD x
// runs the blocks with bid % 8 == x; it gets the contiguous items starting at x * (n / 8) + min(x, n % 8),
// as many as it has blocks, so neighbouring items (image rows, columns) share that XCD's L2.
__device__ __forceinline__ int xcd_row(int bid, int n) {
  const int x = bid % 8, q = n / 8, r = n % 8;
  return x * q + min(x, r) + bid / 8;
}

// ---------------------------------------------------------------- slice extraction
// Element index of (x, y, z) in the template spectrum (x < P/2+1, y and z physical indices < P).
// Layout 0: FFTW order [z][y][x]. Layout 1: bricks of 128 bytes holding 4 (x) x 2 (y) x 2 (z) complex
// values, bricks in [z/2][y/2][x/4] order with the x brick count padded to (P/2+4)/4: a thin slab of
// any orientation touches fewer cache lines than with rows of consecutive x.
template <int LAYOUT>
__device__ __forceinline__ size_t tpl_index(int P, int x, int y, int z) {
  if constexpr (LAYOUT == 0) {
    return ((size_t)z * P + y) * (P / 2 + 1) + x;
  } else {
    const int nxb = (P / 2 + 4) / 4;
    return ((((size_t)(z >> 1) * (P / 2) + (y >> 1)) * nxb + (x >> 2)) << 4) + (((z & 1) << 3) | ((y & 1) << 2) | (x & 3));
  }
}

// Image::ReturnLinearInterpolatedFourier (image.cpp:2899-3020) on a P^3 FFTW half volume.
template <int LAYOUT = 0>
__device__ __forceinline__ cf interp_fourier(const cf* __restrict__ tpl, int P, float x, float y, float z) {
  const int lo = -P / 2, hi_yz = P / 2 - 1, hi_x = P / 2;
  cf sum = cmake(0.f, 0.f);
  const bool pos = x >= 0.0f;
  const int i0 = (int)floorf(x);
  if (pos) { if (i0 + 1 > hi_x) return sum; }
  else { if (i0 < lo) return sum; }
  const int j0 = (int)floorf(y);
  if (j0 < lo || j0 + 1 > hi_yz) return sum;
  const int k0 = (int)floorf(z);
  if (k0 < lo || k0 + 1 > hi_yz) return sum;
#pragma unroll
  for (int dk = 0; dk < 2; ++dk) {
    const int k = k0 + dk;
    const int pz = pos ? (k >= 0 ? k : P + k) : (k > 0 ? P - k : -k);
    const double zw = 1.0 - (double)fabsf(z - (float)k);
#pragma unroll
    for (int dj = 0; dj < 2; ++dj) {
      const int j = j0 + dj;
      const int py = pos ? (j >= 0 ? j : P + j) : (j > 0 ? P - j : -j);
      const double yw = 1.0 - (double)fabsf(y - (float)j);
#pragma unroll
      for (int di = 0; di < 2; ++di) {
        const int i = i0 + di;
        const float w = (float)((1.0 - (double)fabsf(x - (float)i)) * yw * zw);
        cf val = pos ? tpl[tpl_index<LAYOUT>(P, i, py, pz)] : cconj(tpl[tpl_index<LAYOUT>(P, -i, py, pz)]);
        sum = cadd(sum, cscale(val, w));
      }
    }
  }
  return sum;
}

// Slice pixel (i, j) of a P x (P/2+1) slice at orientation matrix m, times the quadrant-swap phase
// and (when filt is not null) the projection filter. Special cases of Image::ExtractSlice: the
// i = 0 column takes Hermitian mates for negative y except the even-size edge row, and the DC term
// is 0 (image.cpp:2556-2589).
template <int LAYOUT = 0>
__device__ __forceinline__ cf slice_value(const cf* __restrict__ tpl, int P, const float* __restrict__ m, int i,
                                          int j, const cf* __restrict__ phase, const cf* __restrict__ filt) {
  const int Ph = P / 2 + 1;
  const int pix = j * Ph + i;
  const int ly = j >= P / 2 ? j - P : j;
  cf val = cmake(0.f, 0.f);
  if (!(i == 0 && ly == 0)) {
    int sy = ly;
    bool mate = false;
    if (i == 0 && ly < 0 && ly != -P / 2) { sy = -ly; mate = true; }
    const float x2 = (float)i, y2 = (float)sy;
    const float x3 = m[0] * x2 + m[1] * y2;
    const float y3 = m[3] * x2 + m[4] * y2;
    const float z3 = m[6] * x2 + m[7] * y2;
    val = interp_fourier<LAYOUT>(tpl, P, x3, y3, z3);
    if (mate) val = cconj(val);
  }
  val = cmul(val, phase[pix]);
  if (filt) val = cmul(val, filt[pix]);
  return val;
}

// Slice extraction fused with the inverse column FFT that starts the slice's C2R (image.cpp:5104-5116):
// a block builds KC adjacent columns of one orientation's slice in LDS and transforms them there,
// so the slice is written once, after the transform, in runs of KC columns. Same per-pixel
// arithmetic and FFT plan (64 threads per column) as extracting the slice and then k_cols<P, 64, +1, 0>.
template <int P, int KC, int LAYOUT = 0>
__global__ __launch_bounds__(KC * 64) void k_extract_cols(const cf* __restrict__ tpl, const float* __restrict__ mats,
                                                         int o0, const cf* __restrict__ phase,
                                                         const cf* __restrict__ filt, const cf* __restrict__ tw,
                                                         cf* __restrict__ out) {
  constexpr int Ph = P / 2 + 1, NTP = 64;
  __shared__ cf buf[KC][lds_len(P)];
  const int b = blockIdx.y, i0 = blockIdx.x * KC, tid = threadIdx.x;
  const float* m = mats + (size_t)(o0 + b) * 9;
  for (int e = tid; e < KC * P; e += KC * NTP) {
    const int c = e % KC, j = e / KC, i = i0 + c;
    buf[c][lpad(j)] = i < Ph ? slice_value<LAYOUT>(tpl, P, m, i, j, phase, filt) : cmake(0.f, 0.f);
  }
  __syncthreads();
  fft_inplace<P, +1, NTP>(buf[tid / NTP], tw, tid % NTP);
  cf* o = out + (size_t)b * P * Ph;
  for (int e = tid; e < KC * P; e += KC * NTP) {
    const int c = e % KC, j = e / KC, i = i0 + c;
    if (i < Ph) o[(size_t)j * Ph + i] = buf[c][lpad(j)];
  }
}

// Copies a P^3 half spectrum from FFTW order into layout 1 (padding bricks entries stay unset and
// are never read: x < P/2+1 always).
static __global__ void k_tpl_to_bricks(const cf* __restrict__ src, int P, cf* __restrict__ dst) {
  const int Ph = P / 2 + 1;
  const size_t n = (size_t)P * P * Ph;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(e % Ph);
    const size_t r = e / Ph;
    const int y = (int)(r % P), z = (int)(r / P);
    dst[tpl_index<1>(P, x, y, z)] = src[e];
  }
}

// ---------------------------------------------------------------- pixel-size change of a volume
// Bluestein transform of any length N <= M/2 along strided lines (one block per line, lines laid out
// as k_cols): X_k = sum_n x_n exp(s 2 pi i n k / N) = c_k (a * b)_k with a_n = x_n c_n, c_n =
// exp(s i pi n^2 / N), b_m = conj(c_|m|) for |m| < N; the circular convolution of length M runs as
// two M-point FFTs. chirp holds c_n (n < N); bhat holds FFT_M(b) / M.
template <int M, int NT>
__global__ __launch_bounds__(NT) void k_bluestein_cols(cf* __restrict__ data, int N, int ncols, size_t bstride,
                                                       const cf* __restrict__ chirp, const cf* __restrict__ bhat,
                                                       const cf* __restrict__ tw) {
  __shared__ cf buf[lds_len(M)];
  const int item = xcd_row(blockIdx.x, gridDim.x), col = item % ncols, b = item / ncols, tid = threadIdx.x;
  cf* base = data + (size_t)b * bstride + col;
  for (int e = tid; e < M; e += NT) buf[lpad(e)] = e < N ? cmul(base[(size_t)e * ncols], chirp[e]) : cmake(0.f, 0.f);
  __syncthreads();
  fft_inplace<M, -1, NT>(buf, tw, tid);
  for (int e = tid; e < M; e += NT) buf[lpad(e)] = cmul(buf[lpad(e)], bhat[e]);
  __syncthreads();
  fft_inplace<M, +1, NT>(buf, tw, tid);
  for (int e = tid; e < N; e += NT) base[(size_t)e * ncols] = cmul(buf[lpad(e)], chirp[e]);
}

// C2R of rows of any even length N <= M/2 (FFTW c2r semantics: the imaginary parts of the DC and
// Nyquist terms are ignored): the Hermitian extension of the N/2+1 inputs, then the inverse Bluestein
// transform above (chirp and bhat for s = +1), real parts out. Rows as k_rows_c2r; in place allowed.
template <int M, int NT>
__global__ __launch_bounds__(NT) void k_bluestein_c2r_rows(const cf* in, int N, size_t ibs, int irs, float* out,
                                                           size_t obs, int ors, const cf* __restrict__ chirp,
                                                           const cf* __restrict__ bhat, const cf* __restrict__ tw) {
  __shared__ cf buf[lds_len(M)];
  const int row = blockIdx.x, b = blockIdx.y, tid = threadIdx.x, h = N / 2;
  const cf* src = in + (size_t)b * ibs + (size_t)row * irs;
  for (int e = tid; e < M; e += NT) {
    cf v = cmake(0.f, 0.f);
    if (e < N) {
      cf X = e <= h ? src[e] : cconj(src[N - e]);
      if (e == 0 || e == h) X.y = 0.f;
      v = cmul(X, chirp[e]);
    }
    buf[lpad(e)] = v;
  }
  __syncthreads();
  fft_inplace<M, -1, NT>(buf, tw, tid);
  for (int e = tid; e < M; e += NT) buf[lpad(e)] = cmul(buf[lpad(e)], bhat[e]);
  __syncthreads();
  fft_inplace<M, +1, NT>(buf, tw, tid);
  float* dst = out + (size_t)b * obs + (size_t)row * ors;
  for (int e = tid; e < N; e += NT) dst[e] = cmul(buf[lpad(e)], chirp[e]).x;
}

// x *= k over n complex values (the ForwardFFT scale of a large spectrum).
static __global__ void k_scale_cf(cf* __restrict__ x, size_t n, float k) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    x[e] = cmake(x[e].x * k, x[e].y * k);
}

// Image::ClipInto, real space, centred (image.cpp:8556-8586), between cubes in the padded layout:
// out (edge b) from in (edge a), padding value outside.
static __global__ void k_real_clip3(const float* __restrict__ in, int a, float* __restrict__ out, int b, float pad) {
  const size_t n = (size_t)b * b * (b + 2);
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int ii = (int)(e % (b + 2));
    const size_t r = e / (b + 2);
    const int jj = (int)(r % b), kk = (int)(r / b);
    float v = 0.0f;
    if (ii < b) {
      const int i = a / 2 + (ii - b / 2), j = a / 2 + (jj - b / 2), k = a / 2 + (kk - b / 2);
      v = (i < 0 || i >= a || j < 0 || j >= a || k < 0 || k >= a) ? pad : in[((size_t)k * a + j) * (a + 2) + i];
    }
    out[e] = v;
  }
}

// Image::ClipInto, Fourier space (image.cpp:8588-8620), between cubic half spectra (edges a -> b,
// both even): out's physical (ii, jj, kk) takes in's value at the same logical coordinate, 0 outside
// in's logical bounds.
static __global__ void k_fourier_clip3(const cf* __restrict__ in, int a, cf* __restrict__ out, int b) {
  const int ah = a / 2 + 1, bh = b / 2 + 1;
  const size_t n = (size_t)bh * b * b;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int ii = (int)(e % bh);
    const size_t r = e / bh;
    const int jj = (int)(r % b), kk = (int)(r / b);
    const int ly = jj >= b / 2 ? jj - b : jj, lz = kk >= b / 2 ? kk - b : kk;
    cf v = cmake(0.f, 0.f);
    if (ii <= a / 2 && ly >= -a / 2 && ly <= a / 2 - 1 && lz >= -a / 2 && lz <= a / 2 - 1) {
      const int py = ly >= 0 ? ly : a + ly, pz = lz >= 0 ? lz : a + lz;
      v = in[((size_t)pz * a + py) * ah + ii];
    }
    out[e] = v;
  }
}

// The Nyquist planes ClipInto fills in when a half spectrum grows (image.cpp:8626-8690), step 1: the
// positive y Nyquist row of every logical z of the smaller cube takes the negative one.
static __global__ void k_fourier_grow_y(const cf* __restrict__ in, int a, cf* __restrict__ out, int b) {
  const int ah = a / 2 + 1, bh = b / 2 + 1;
  const int n = ah * a;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const int ii = e % ah, kl = e / ah - a / 2;
    const int pzo = kl >= 0 ? kl : b + kl, pzi = kl >= 0 ? kl : a + kl;
    out[((size_t)pzo * b + a / 2) * bh + ii] = in[((size_t)pzi * a + a / 2) * ah + ii];
  }
}

// Step 2: the positive z Nyquist plane (physical a/2 of the larger cube) takes the rows of the
// negative one (physical b - a/2) for physical y in 0..a/2 and b-a/2..b-1.
static __global__ void k_fourier_grow_z(cf* __restrict__ out, int a, int b) {
  const int ah = a / 2 + 1, bh = b / 2 + 1;
  const int rows = a + 1;  // jj = 0..a/2 and b-a/2..b-1
  const int n = ah * rows;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const int ii = e % ah, q = e / ah;
    const int jj = q <= a / 2 ? q : b - a / 2 + (q - a / 2 - 1);
    out[((size_t)(a / 2) * b + jj) * bh + ii] = out[((size_t)(b - a / 2) * b + jj) * bh + ii];
  }
}

// ---------------------------------------------------------------- pixel-size change of projections
// The same steps for a batch of square images (one per item, edges a -> b): masked mean inside
// 0.45 a (image.cpp:6263-6322) per item, real clips with a per-item padding value, Fourier clips
// with the 2-D Nyquist row ClipInto fills in when an image grows (image.cpp:8628-8635).
template <int NT>
__global__ __launch_bounds__(NT) void k_masked_mean2(const float* __restrict__ in, int a, size_t ibs, float radius,
                                                     float* __restrict__ out) {
  __shared__ double rs[NT];
  __shared__ long long rn[NT];
  const float* im = in + (size_t)blockIdx.x * ibs;
  const float r2 = radius * radius;
  double sum = 0.0;
  long long n = 0;
  for (int e = threadIdx.x; e < a * a; e += NT) {
    const int j = e / a, i = e - j * a;
    const float dy = (float)(j - a / 2), dx = (float)(i - a / 2);
    const float d2 = (float)((double)dx * dx) + (float)((double)dy * dy) + 0.0f;
    if (d2 <= r2) { sum += im[(size_t)j * (a + 2) + i]; ++n; }
  }
  rs[threadIdx.x] = sum;
  rn[threadIdx.x] = n;
  __syncthreads();
  for (int st = NT / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) { rs[threadIdx.x] += rs[threadIdx.x + st]; rn[threadIdx.x] += rn[threadIdx.x + st]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = rn[0] > 0 ? (float)(rs[0] / rn[0]) : 0.0f;
}

static __global__ void k_real_clip2(const float* __restrict__ in, int a, size_t ibs, float* __restrict__ out, int b,
                                    size_t obs, const float* __restrict__ pad) {
  const int item = blockIdx.y;
  const float pv = pad[item];
  const float* im = in + (size_t)item * ibs;
  float* om = out + (size_t)item * obs;
  const int n = b * (b + 2);
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const int ii = e % (b + 2), jj = e / (b + 2);
    float v = 0.0f;
    if (ii < b) {
      const int i = a / 2 + (ii - b / 2), j = a / 2 + (jj - b / 2);
      v = (i < 0 || i >= a || j < 0 || j >= a) ? pv : im[(size_t)j * (a + 2) + i];
    }
    om[e] = v;
  }
}

static __global__ void k_fourier_clip2(const cf* __restrict__ in, int a, size_t ibs, cf* __restrict__ out, int b,
                                       size_t obs) {
  const int item = blockIdx.y, ah = a / 2 + 1, bh = b / 2 + 1;
  const cf* im = in + (size_t)item * ibs;
  cf* om = out + (size_t)item * obs;
  const int n = bh * b;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const int ii = e % bh, jj = e / bh;
    const int ly = jj >= b / 2 ? jj - b : jj;
    cf v = cmake(0.f, 0.f);
    if (ii <= a / 2 && ly >= -a / 2 && ly <= a / 2 - 1) v = im[(size_t)(ly >= 0 ? ly : a + ly) * ah + ii];
    if (b > a && jj == a / 2 && ii <= a / 2) v = im[(size_t)(a / 2) * ah + ii];  // the positive Nyquist row
    om[e] = v;
  }
}

// The spectra of the t x t projections scaled by k (ForwardFFT), copied to their batch slots.
static __global__ void k_copy_scale_cf(const cf* __restrict__ in, size_t ibs, cf* __restrict__ out, size_t obs, int n,
                                       float k) {
  const int item = blockIdx.y;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const cf v = in[(size_t)item * ibs + e];
    out[(size_t)item * obs + e] = cmake(v.x * k, v.y * k);
  }
}

// ---------------------------------------------------------------- small line transforms
// Complex FFTs of length N down columns of [b][N][ncols] complex arrays (row stride ncols).
// MODE 0: one transform of sign SIGN. MODE 1: forward, multiply by scale * filt[row][col], inverse
// (ForwardFFT scale, MultiplyPixelWise, first half of BackwardFFT; match_template.cpp:1019-1030).
// Launched as a one-dimensional grid of ncols * nb blocks.
template <int N, int NT, int SIGN, int MODE>
__global__ __launch_bounds__(NT) void k_cols(cf* __restrict__ data, int ncols, size_t bstride, const cf* __restrict__ tw,
                                             const cf* __restrict__ filt, float scale) {
  __shared__ cf buf[lds_len(N)];
  // one-dimensional grid of ncols x nb blocks; neighbouring columns run on one XCD and share the
  // cache lines of their rows there
  const int item = xcd_row(blockIdx.x, gridDim.x), col = item % ncols, b = item / ncols, tid = threadIdx.x;
  cf* base = data + (size_t)b * bstride + col;
  for (int e = tid; e < N; e += NT) buf[lpad(e)] = base[(size_t)e * ncols];
  __syncthreads();
  if constexpr (MODE == 0) {
    fft_inplace<N, SIGN, NT>(buf, tw, tid);
  } else {
    fft_inplace<N, -1, NT>(buf, tw, tid);
    for (int e = tid; e < N; e += NT) buf[lpad(e)] = cmul(cscale(buf[lpad(e)], scale), filt[(size_t)e * ncols + col]);
    __syncthreads();
    fft_inplace<N, +1, NT>(buf, tw, tid);
  }
  for (int e = tid; e < N; e += NT) base[(size_t)e * ncols] = buf[lpad(e)];
}

// C2R of rows of length N (FFTW c2r semantics, unscaled) via an N/2 complex inverse FFT.
// Input row: N/2+1 complex at in + b*ibs + row*irs (complex units); output N reals at
// out + b*obs + row*ors (float units). In place is allowed (all reads precede writes).
template <int N, int NT>
__global__ __launch_bounds__(NT) void k_rows_c2r(const cf* in, size_t ibs, int irs, float* out, size_t obs, int ors,
                                                 const cf* __restrict__ tw_n, const cf* __restrict__ tw_h) {
  constexpr int h = N / 2;
  __shared__ cf buf[lds_len(h + 1)];
  const int row = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
  const cf* src = in + (size_t)b * ibs + (size_t)row * irs;
  for (int e = tid; e <= h; e += NT) buf[lpad(e)] = src[e];
  __syncthreads();
  // Z_k = (X_k + conj X_{h-k}) + i (X_k - conj X_{h-k}) conj(W_N^k), pairs (k, h-k)
  for (int k = tid; k <= h / 2; k += NT) {
    if (k == 0) {
      const float x0 = buf[lpad(0)].x, xh = buf[lpad(h)].x;
      buf[lpad(0)] = cmake(x0 + xh, x0 - xh);
    } else {
      const int kk = h - k;
      const cf Xk = buf[lpad(k)], Xc = cconj(buf[lpad(kk)]);
      const cf E = cadd(Xk, Xc);
      const cf D = csub(Xk, Xc);
      const cf w = tw_n[k];
      const cf O = cmul(D, cconj(w));
      const cf Ok = cmul(cconj(D), w);
      const cf Zk = cmake(E.x - O.y, E.y + O.x);
      const cf Zkk = cmake(E.x - Ok.y, -E.y + Ok.x);
      buf[lpad(k)] = Zk;
      buf[lpad(kk)] = Zkk;
    }
  }
  __syncthreads();
  fft_inplace<h, +1, NT>(buf, tw_h, tid);
  float* dst = out + (size_t)b * obs + (size_t)row * ors;
  for (int m = tid; m < h; m += NT) {
    const cf z = buf[lpad(m)];
    dst[2 * m] = z.x;
    dst[2 * m + 1] = z.y;
  }
}

// ---------------------------------------------------------------- inter-pass plane storage
// Storage of the plane Z between the column and row transforms: ZS_F32 (FP32 complex, 8 bytes),
// ZS_F16 or ZS_BF16 (two 16-bit halves packed in 4 bytes, real in the low half). 16-bit values are
// stored multiplied by a power of two (exact) that brings the plane's typical magnitude near 1, so
// FP16 keeps its full precision; the row pass multiplies by the inverse after loading.
enum { ZS_F32 = 0, ZS_F16 = 1, ZS_BF16 = 2 };

template <int ZST>
__device__ __forceinline__ unsigned short to_half_bits(float x) {
  if constexpr (ZST == ZS_F16) {
    const _Float16 h = (_Float16)x;  // round to nearest even
    unsigned short b;
    __builtin_memcpy(&b, &h, sizeof(b));
    return b;
  } else {
    // bfloat16, round to nearest even (finite values)
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
  }
}
template <int ZST>
__device__ __forceinline__ float from_half_bits(unsigned b) {
  if constexpr (ZST == ZS_F16) {
    const unsigned short s = (unsigned short)b;
    _Float16 h;
    __builtin_memcpy(&h, &s, sizeof(h));
    return (float)h;
  } else {
    return __uint_as_float(b << 16);
  }
}
template <int ZST>
__device__ __forceinline__ unsigned pack_half2(cf v, float scale) {
  return (unsigned)to_half_bits<ZST>(v.x * scale) | ((unsigned)to_half_bits<ZST>(v.y * scale) << 16);
}

// R2C of rows of length N (unscaled, then * scale) via an N/2 complex forward FFT.
// The input row is zero except a window of `len` reals starting at column c0 of the N-row,
// read from src + b*sbs + (r0 + row)*srs + s0 (float units). Output N/2+1 complex at
// dst + b*dbs + row*drs (element units): FP32 complex, or with OS = ZS_F16 / ZS_BF16 one 32-bit
// word per element holding the two 16-bit parts of the scaled value (pack_half2).
template <int N, int NT, int OS = ZS_F32>
__global__ __launch_bounds__(NT) void k_rows_r2c(const float* __restrict__ src, size_t sbs, int srs, int r0, int s0,
                                                 int len, int c0, cf* __restrict__ dst, size_t dbs, int drs,
                                                 const cf* __restrict__ tw_n, const cf* __restrict__ tw_h,
                                                 float scale) {
  constexpr int h = N / 2;
  __shared__ cf buf[lds_len(h + 1)];
  const int row = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
  const float* s = src + (size_t)b * sbs + (size_t)(r0 + row) * srs + s0;
  for (int m = tid; m < h; m += NT) {
    const int x0 = 2 * m - c0, x1 = 2 * m + 1 - c0;
    const float a = (x0 >= 0 && x0 < len) ? s[x0] : 0.f;
    const float c = (x1 >= 0 && x1 < len) ? s[x1] : 0.f;
    buf[lpad(m)] = cmake(a, c);
  }
  __syncthreads();
  fft_inplace<h, -1, NT>(buf, tw_h, tid);
  // X_k = E_k + W^k O_k, X_{h-k} = conj(E_k - W^k O_k); E = (Z_k + conj Z_{h-k})/2, O = (Z_k - conj Z_{h-k})/(2i)
  const size_t d0 = (size_t)b * dbs + (size_t)row * drs;
  auto put = [&](int k, cf v) {  // element k of the row: v * scale
    if constexpr (OS == ZS_F32) dst[d0 + k] = cscale(v, scale);
    else reinterpret_cast<unsigned*>(dst)[d0 + k] = pack_half2<OS>(v, scale);
  };
  for (int k = tid; k <= h / 2; k += NT) {
    if (k == 0) {
      const cf z0 = buf[lpad(0)];
      put(0, cmake(z0.x + z0.y, 0.f));
      put(h, cmake(z0.x - z0.y, 0.f));
    } else {
      const int kk = h - k;
      const cf zk = buf[lpad(k)], zc = cconj(buf[lpad(kk)]);
      const cf E = cscale(cadd(zk, zc), 0.5f);
      const cf D = cscale(csub(zk, zc), 0.5f);
      const cf O = cmake(D.y, -D.x);
      const cf wO = cmul(tw_n[k], O);
      const cf Xk = cadd(E, wO);
      const cf Xkk = cconj(csub(E, wO));
      put(k, Xk);
      put(kk, Xkk);
    }
  }
}

// ---------------------------------------------------------------- projection normalisation
// A T x T real image in the padded layout (row stride T+2), normalised as match_template.cpp:1060-1068:
// subtract the edge mean (image.cpp:5907-6000), then variance = <x^2> * nT/nP - (<x> * nT/nP)^2 in
// float from FP64 sums, then x * (1/sqrt(var)). S blocks per orientation (blockIdx.x = band of T/S
// rows, blockIdx.y = orientation): k_norm_stats forms the edge mean (every block the same FP64 tree
// sum) and its band's FP64 sums of x - mean without writing the image (other blocks still read its
// edges); k_norm_scale adds the S band sums in band order and writes (x - mean) * (1/sqrt(var)).
// part[b][s] = {sum, sum of squares, edge mean}.
template <int NT>
__device__ void norm_band(int T, int s, int S, int& lo, int& n, int& sh) {
  const int y0 = T * s / S, y1 = T * (s + 1) / S;
  lo = y0;
  n = (y1 - y0) * T;
  sh = (T & (T - 1)) == 0 ? __ffs(T) - 1 : -1;  // element i of the band: row i >> sh, column i & (T - 1)
}
template <int NT>
__global__ __launch_bounds__(NT) void k_norm_stats(const float* __restrict__ proj, int T, size_t bstride,
                                                   double* __restrict__ part) {
  __shared__ double red[2][NT];
  const int tid = threadIdx.x, s = blockIdx.x, S = gridDim.x, b = blockIdx.y;
  const float* p = proj + (size_t)b * bstride;
  const int rs = T + 2;
  double e = 0.0;
  for (int i = tid; i < 4 * T - 4; i += NT) {
    int x, y;
    if (i < T) { x = i; y = 0; }
    else if (i < 2 * T) { x = i - T; y = T - 1; }
    else { const int r = i - 2 * T; y = 1 + r / 2; x = (r & 1) ? T - 1 : 0; }
    e += (double)p[y * rs + x];
  }
  red[0][tid] = e;
  __syncthreads();
  for (int k = NT / 2; k > 0; k >>= 1) {
    if (tid < k) red[0][tid] += red[0][tid + k];
    __syncthreads();
  }
  const float edge_mean = (float)(red[0][0] / (double)(float)(4 * T - 4));
  __syncthreads();
  int y0, n, sh;
  norm_band<NT>(T, s, S, y0, n, sh);
  double s1 = 0.0, s2 = 0.0;
  for (int i = tid; i < n; i += NT) {
    const int dy = sh >= 0 ? i >> sh : i / T, x = i - dy * T;
    const float v = p[(y0 + dy) * rs + x] + (-edge_mean);
    s1 += (double)v;
    s2 += (double)v * (double)v;
  }
  red[0][tid] = s1;
  red[1][tid] = s2;
  __syncthreads();
  for (int k = NT / 2; k > 0; k >>= 1) {
    if (tid < k) { red[0][tid] += red[0][tid + k]; red[1][tid] += red[1][tid + k]; }
    __syncthreads();
  }
  if (tid == 0) {
    double* q = part + ((size_t)b * S + s) * 3;
    q[0] = red[0][0];
    q[1] = red[1][0];
    q[2] = (double)edge_mean;
  }
}
template <int NT>
__global__ __launch_bounds__(NT) void k_norm_scale(float* __restrict__ proj, int T, size_t bstride,
                                                   const double* __restrict__ part, long long n_pad) {
  __shared__ float mi[2];
  const int tid = threadIdx.x, s = blockIdx.x, S = gridDim.x, b = blockIdx.y;
  if (tid == 0) {
    const double* q = part + (size_t)b * S * 3;
    double s1 = 0.0, s2 = 0.0;
    for (int k = 0; k < S; ++k) {
      s1 += q[3 * k];
      s2 += q[3 * k + 1];
    }
    const float mean_sq = (float)(s2 / T / T);                        // ReturnSumOfSquares (image.cpp:294-301)
    const float mean = (float)(s1 / ((long long)T * T));              // ReturnAverageOfRealValues (image.cpp:6327-6340)
    const float nT = (float)((long long)T * T), nP = (float)n_pad;
    const float t1 = mean_sq * nT / nP;
    const float t2 = mean * nT / nP;
    const float variance = t1 - t2 * t2;
    mi[0] = (float)q[2];
    mi[1] = (float)(1. / (double)sqrtf(variance));                   // DivideByConstant (image.cpp:5118-5127)
  }
  __syncthreads();
  const float edge_mean = mi[0], inverse = mi[1];
  float* p = proj + (size_t)b * bstride;
  const int rs = T + 2;
  int y0, n, sh;
  norm_band<NT>(T, s, S, y0, n, sh);
  for (int i = tid; i < n; i += NT) {
    const int dy = sh >= 0 ? i >> sh : i / T, x = i - dy * T;
    float* e = p + (y0 + dy) * rs + x;
    const float v = *e + (-edge_mean);
    *e = v * inverse;
  }
}

template <int ZST>
__device__ __forceinline__ cf unpack_half2(unsigned w, float inv_scale) {
  return cmake(from_half_bits<ZST>(w & 0xFFFFu) * inv_scale, from_half_bits<ZST>(w >> 16) * inv_scale);
}
// Element i of the forward rows, stored like the plane: FP32 complex, or 16-bit pairs scaled by the
// plane's power-of-two scale (k_rows_r2c<.., ZST>), widened and unscaled here (exact).
template <int ZST>
__device__ __forceinline__ cf row_elem(const cf* __restrict__ rows, size_t i, float inv_scale) {
  if constexpr (ZST == ZS_F32) return rows[i];
  else return unpack_half2<ZST>(reinterpret_cast<const unsigned*>(rows)[i], inv_scale);
}

// ---------------------------------------------------------------- column pass
// Tile of KC adjacent columns kx of the W x H spectrum, NTC threads per column. For each
// orientation b: place the T forward-row-FFT rows at y = H/2 - T/2 + r, forward column FFT,
// zero the DC term (ZeroCentralPixel, image.cpp:6498-6503), multiply conj(R) * I
// (match_template.cpp:1083), inverse column FFT, and write Z[b][tile][y][KC] (column tiles
// contiguous, so every store of the tile is a full cache line).
// Tiles are mapped so that consecutive tiles run on the same XCD (workgroups are dispatched
// round-robin over the 8 XCDs), letting partial cache lines of Z merge in that XCD's L2. The
// orientations of the batch are spread over gridDim.y (block row y takes b = y, y + gridDim.y, ..),
// so that small images, with few column tiles, still fill the GPU. Each butterfly reads one twiddle
// from the global table and forms the others of its stage as powers of it.
// Largest power-of-two radix of the column pass's plan (the remaining factors of two form one smaller
// stage): 2592 = 2^5 3^4 runs 8 4 9 9 (324, 648, 288, 288 butterflies for its 324 threads per column)
// rather than 16 2 9 9 (162 butterflies in the radix-16 stage).
constexpr int mid_lr(int H) { return H == 2592 ? 3 : 4; }
template <int H, int KC, int NTC, int ZST = ZS_F32>
__global__ __launch_bounds__(KC* NTC) void k_mid(const cf* __restrict__ rows, const cf* __restrict__ inpT,
                                                 cf* __restrict__ Z, int Wh, int T, int nb, int ntiles,
                                                 const cf* __restrict__ tw, float zscale) {
  __shared__ cf buf[KC][lds_len(H)];
  const int tid = threadIdx.x;
  const int c = tid / NTC, ctid = tid % NTC;
  const int bid = blockIdx.x;
  const int per_xcd = (ntiles + 7) / 8;
  const int tile = (bid % 8) * per_xcd + bid / 8;
  if (tile >= ntiles) return;  // whole block exits together
  const int kx0 = tile * KC;
  const int kx = kx0 + c;
  const bool valid = kx < Wh;
  const int jlo = H / 2 - T / 2;
  const size_t rows_bs = (size_t)T * Wh;
  const size_t z_bs = (size_t)ntiles * H * KC;
  const float rinv = 1.0f / zscale;  // a power of two: exact
  for (int b = blockIdx.y; b < nb; b += gridDim.y) {
    for (int e = ctid; e < H; e += NTC) buf[c][lpad(e)] = cmake(0.f, 0.f);
    __syncthreads();
    const size_t rb = b * rows_bs;
    for (int e = tid; e < KC * T; e += KC * NTC) {
      const int cc = e % KC, r = e / KC;
      const int kxx = kx0 + cc;
      if (kxx < Wh) buf[cc][lpad(jlo + r)] = row_elem<ZST>(rows, rb + (size_t)r * Wh + kxx, rinv);
    }
    __syncthreads();
    fft_inplace<H, -1, NTC, mid_lr(H)>(buf[c], TwPow{tw}, ctid);
    if (valid) {
      const cf* ic = inpT + (size_t)kx * H;
      for (int e = ctid; e < H; e += NTC) {
        cf v = buf[c][lpad(e)];
        if (kx == 0 && e == 0) v = cmake(0.f, 0.f);
        buf[c][lpad(e)] = cconjmul(v, ic[e]);
      }
    }
    __syncthreads();
    fft_inplace<H, +1, NTC, mid_lr(H)>(buf[c], TwPow{tw}, ctid);
    // Z[b][tile][y][KC]: the tile's output is one contiguous block; streamed past L2
    const size_t zo = b * z_bs + (size_t)tile * H * KC;
    for (int e = tid; e < KC * H; e += KC * NTC) {
      const int cc = e % KC, y = e / KC;
      const cf v = buf[cc][lpad(y)];
      if constexpr (ZST == ZS_F32) {
        unsigned long long bits;
        __builtin_memcpy(&bits, &v, sizeof(bits));
        __builtin_nontemporal_store(bits, reinterpret_cast<unsigned long long*>(Z + zo + e));
      } else {
        __builtin_nontemporal_store(pack_half2<ZST>(v, zscale), reinterpret_cast<unsigned*>(Z) + zo + e);
      }
    }
    __syncthreads();
  }
}

// Column pass for H = 4096 with the data register-resident between the three radix-16 stages:
// KC * 256 threads = 256 butterflies x KC columns, lane -> (column c = tid % KC, butterfly j = tid / KC).
// With KC = 2 a block needs ~74 KB of LDS, so two blocks share a CU and hide each other's barrier
// and load latency.
//  * forward stage 1 reads its 16 inputs x[j + 256 q] straight from the forward rows (zero outside
//    the T projection rows), so no LDS zero-fill or scatter is needed; with DIRECT (T <= 256) the
//    stage is trivial and stage 2 loads its inputs from the rows instead (one exchange fewer);
//  * forward stage 3, the DC zero, conj(R) * I and inverse stage 1 run on registers;
//  * inverse stage 3 leaves z[j + 256 q] in registers; column kx goes to slot kx % ZT of tile
//    kx / ZT of Z[b][tile][y][ZT] (ZT a multiple of KC: with ZT = 2 KC two neighbouring blocks,
//    which run on the same XCD, fill one tile, and a row read of k_last takes ZT * 8 contiguous
//    bytes per tile); NTS selects non-temporal stores;
//  * barriers order LDS traffic only, so they do not wait for the Z stores to complete.
// Eight LDS passes per column and orientation (four exchanges, each a write and a read). With the
// padded LDS index lpad(i) = i + i/16 the three access patterns are a per-thread base plus
// compile-time offsets: lpad(16 j + q) = 17 j + q, lpad(j + 256 q) = lpad(j) + 272 q and
// lpad(b2 + 16 q) = lpad(b2) + 17 q for b2 = (j / 16) 256 + j % 16.
template <int KC, bool PREFETCH, int ZT = KC, bool NTS = true, bool DIRECT = false, int ZST = ZS_F32>
__global__ __launch_bounds__(KC * 256, 4) void k_mid4096(const cf* __restrict__ rows, const cf* __restrict__ inpT,
                                                  cf* __restrict__ Z, int Wh, int T, int nb, int ntiles,
                                                  const cf* __restrict__ tw, float zscale) {
  constexpr int H = 4096, M = 256;
  constexpr int CS = lds_len(H) + 8;  // column stride: shifts the columns by 16 banks
  static_assert(KC * M <= 1024, "KC columns of 256 butterflies per block");
  const float rinv = 1.0f / zscale;  // 16-bit forward rows carry the plane's power-of-two scale
  __shared__ cf lds[KC * CS];
  // Twiddles from LDS: W4096^(16 m) = W256^m; the stage-2 twiddle W256^(k q) and the second factor
  // of the stage-3 twiddle W4096^(j q) = W256^(a q) * W4096^(b q) (j = 16 a + b) are stored q-major,
  // [q][k] and [q][b], so that the 16 lanes of a wave that differ in k (or b) read 16 consecutive
  // entries (distinct banks) rather than entries 16 apart (one bank).
  __shared__ cf tw256[256];
  __shared__ cf tw2q[256];
  __shared__ cf tw3q[256];
  const int tid = threadIdx.x;
  if (tid < 256) {
    const int q = tid / 16, k = tid % 16;
    tw256[tid] = tw[16 * tid];
    tw2q[tid] = tw[16 * ((k * q) % 256)];
    tw3q[tid] = tw[k * q];
  }
  lds_barrier();
  const int c = tid % KC;
  const int bid = blockIdx.x;
  const int per_xcd = (ntiles + 7) / 8;
  const int tile = (bid % 8) * per_xcd + bid / 8;
  if (tile >= ntiles) return;  // whole block exits together
  const int kx = tile * KC + c;
  const bool valid = kx < Wh;
  // The T projection rows sit at y = 0..T-1 of the column; the reference centres them at
  // y = H/2 - T/2 (ClipIntoLargerRealSpace2D), a shift whose phase e^{2 pi i (H/2 - T/2) ky / H} the
  // caller has folded into the input spectrum inpT.
  static_assert(ZT % KC == 0, "Z tiles hold whole column groups");
  const size_t z_bs = (size_t)((Wh + ZT - 1) / ZT) * H * ZT;
  for (int b = 0; b < nb; ++b) {
    // Indices are re-derived every iteration from a register the compiler cannot see through, so
    // the address and twiddle computations are not hoisted out of the loop (they would spill).
    int j = tid / KC;
    asm volatile("" : "+v"(j));
    const int k2 = j % 16;
    cf* const colbuf = lds + c * CS;
    cf* const p_s1 = colbuf + 17 * j;                         // stage-1 stores: + q
    cf* const p_rd = colbuf + j + (j >> 4);                   // strided reads:  + 272 q
    cf* const p_s2 = colbuf + (j / 16) * 272 + k2;            // stage-2 stores: + 17 q
    const cf* ic = inpT + (size_t)(valid ? kx : 0) * H + j;   // input spectrum: + 256 q
    cf v[16];
    const size_t rb = (size_t)b * T * Wh + kx;
    if constexpr (DIRECT) {
      // T <= 256: forward stage 1 (Ns = 1) has the single non-zero input x[j] (q = 0), so its 16
      // outputs all equal x[j] and element 16 j' + k1 of its result is x[j']. Stage 2 butterfly j
      // reads elements j + 256 q, which are x[j / 16 + 16 q]: load them from the rows directly.
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int r = j / 16 + 16 * q;
        v[q] = (valid && r < T) ? row_elem<ZST>(rows, rb + (size_t)r * Wh, rinv) : cmake(0.f, 0.f);
      }
    } else {
      // forward stage 1 (Ns = 1): inputs x[j + 256 q] from the T non-zero rows; only q <= (T-1)/256
      // can be non-zero (a block-uniform range)
      const int qmax = (T - 1) / M;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        v[q] = cmake(0.f, 0.f);
        if (q <= qmax) {
          const int r = j + M * q;
          const size_t rq = rb + (size_t)(M * q) * Wh;  // block-uniform base
          if (valid && r < T) v[q] = row_elem<ZST>(rows, rq + (size_t)j * Wh, rinv);
        }
      }
      dft16<-1>(v);
#pragma unroll
      for (int q = 0; q < 16; ++q) p_s1[q] = v[q];
      lds_barrier();
      // forward stage 2 (Ns = 16)
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = p_rd[272 * q];
      lds_barrier();
    }
#pragma unroll
    for (int q = 1; q < 16; ++q) v[q] = cmul(v[q], twiddle<-1>(tw2q, 16 * q + k2));
    dft16<-1>(v);
#pragma unroll
    for (int q = 0; q < 16; ++q) p_s2[17 * q] = v[q];
    lds_barrier();
    // input-spectrum column, optionally prefetched before forward stage 3 so that its latency
    // overlaps the stage (costs 32 registers)
    cf iv[16];
    if constexpr (PREFETCH) {
#pragma unroll
      for (int q = 0; q < 16; ++q) iv[q] = valid ? ic[M * q] : cmake(0.f, 0.f);
    }
    // forward stage 3 (Ns = 256): X[j + 256 q] in registers
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = p_rd[272 * q];
#pragma unroll
    for (int q = 1; q < 16; ++q)
      v[q] = cmul(v[q], twiddle<-1>(tw256, (j / 16) * q), twiddle<-1>(tw3q, 16 * q + k2));
    dft16<-1>(v);
    // ZeroCentralPixel and conj(reference) * input
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      if constexpr (!PREFETCH) iv[q] = valid ? ic[M * q] : cmake(0.f, 0.f);
      const cf x = (kx == 0 && j == 0 && q == 0) ? cmake(0.f, 0.f) : v[q];
      v[q] = cconjmul(x, iv[q]);
    }
    lds_barrier();
    // inverse stage 1
    dft16<+1>(v);
#pragma unroll
    for (int q = 0; q < 16; ++q) p_s1[q] = v[q];
    lds_barrier();
    // inverse stage 2
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = p_rd[272 * q];
    lds_barrier();
#pragma unroll
    for (int q = 1; q < 16; ++q) v[q] = cmul(v[q], twiddle<+1>(tw2q, 16 * q + k2));
    dft16<+1>(v);
#pragma unroll
    for (int q = 0; q < 16; ++q) p_s2[17 * q] = v[q];
    lds_barrier();
    // inverse stage 3: z[j + 256 q]
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = p_rd[272 * q];
#pragma unroll
    for (int q = 1; q < 16; ++q)
      v[q] = cmul(v[q], twiddle<+1>(tw256, (j / 16) * q), twiddle<+1>(tw3q, 16 * q + k2));
    dft16<+1>(v);
    const size_t zoff = b * z_bs + (size_t)(kx / ZT) * H * ZT + (size_t)j * ZT + kx % ZT;  // elements
    if constexpr (ZST == ZS_F32) {
      unsigned long long* zt = reinterpret_cast<unsigned long long*>(Z) + zoff;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        unsigned long long bits;
        __builtin_memcpy(&bits, &v[q], sizeof(bits));
        if constexpr (NTS) __builtin_nontemporal_store(bits, zt + (size_t)M * ZT * q);
        else zt[(size_t)M * ZT * q] = bits;
      }
    } else {
      unsigned* zt = reinterpret_cast<unsigned*>(Z) + zoff;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const unsigned w = pack_half2<ZST>(v[q], zscale);
        if constexpr (NTS) __builtin_nontemporal_store(w, zt + (size_t)M * ZT * q);
        else zt[(size_t)M * ZT * q] = w;
      }
    }
    lds_barrier();  // LDS reads of this orientation complete before the next one writes
  }
}

// Persistent column pass for H = 4096 and T <= 256 (the arithmetic of k_mid4096<2, *, *, *, true>,
// bit for bit). The grid is the number of co-resident blocks (two per CU); the work is a list of
// units (orientation b, pair of columns) in orientation-major order, cut into eight contiguous
// ranges, one per XCD (workgroups are dispatched round-robin over the XCDs, so block bid runs on
// XCD bid % 8), which its blocks take in turn. The blocks of an XCD therefore work on neighbouring
// column pairs of one orientation at a time: they share that orientation's forward rows in L2,
// and a unit's Z writes never wait on a slow block. Three orderings keep memory latency off the
// critical path:
//  * Z uses tiles of two columns (ZT = 2), so one wavefront's 16 stores each write 32 rows x 16 B,
//    whole 128-byte lines, with no partner block to merge with;
//  * the input-spectrum column is loaded at the top of the unit, beside the rows, and its latency
//    hides behind forward stages 2 and 3;
//  * a unit's Z stores are issued after the next unit's row and input-spectrum loads (ROT): vmcnt
//    counts loads and stores in issue order, so a wait for loads issued after the stores would also
//    wait for the stores to be acknowledged.
// With ZT = 4 (the k_mid4096 layout) the two column pairs of a tile are neighbouring units, run at
// the same time by two blocks of one XCD; their half rows are then stored through L2 (not
// non-temporal), where they merge into whole lines, and are issued in place (ROT off).
template <int ZST = ZS_F32, int ZT = 2>
__global__ __launch_bounds__(512, 4) void k_mid4096p(const cf* __restrict__ rows, const cf* __restrict__ inpT,
                                                     cf* __restrict__ Z, int Wh, int T, int nb, int npairs,
                                                     const cf* __restrict__ tw, float zscale) {
  constexpr int H = 4096, M = 256, KC = 2;
  constexpr bool NTS = ZT == 2, ROT = ZT == 2;
  const float rinv = 1.0f / zscale;  // 16-bit forward rows carry the plane's power-of-two scale
  static_assert(ZT == 2 || ZT == 4, "Z tiles of 2 or 4 columns");
  constexpr int CS = lds_len(H) + 8;  // column stride: shifts the columns by 16 banks
  __shared__ cf lds[KC * CS];
  __shared__ cf tw256[256];
  __shared__ cf tw2q[256];
  __shared__ cf tw3q[256];
  const int tid = threadIdx.x;
  if (tid < 256) {
    const int q = tid / 16, k = tid % 16;
    tw256[tid] = tw[16 * tid];
    tw2q[tid] = tw[16 * ((k * q) % 256)];
    tw3q[tid] = tw[k * q];
  }
  lds_barrier();
  const int c = tid % KC;
  const long long U = (long long)npairs * nb;
  const int x = blockIdx.x % 8, nbx = gridDim.x / 8, l = blockIdx.x / 8;
  const long long u_lo = U * x / 8, u_hi = U * (x + 1) / 8;
  const size_t z_bs = (size_t)((Wh + ZT - 1) / ZT) * H * ZT;
  cf zp[16];  // the previous unit's Z column segment, stored once the next unit's loads are issued
  size_t zoff_prev = 0;
  bool have_prev = false;
  auto store_z = [&](const cf* val, size_t zoff) {
    if constexpr (ZST == ZS_F32) {
      unsigned long long* zt = reinterpret_cast<unsigned long long*>(Z) + zoff;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        unsigned long long bits;
        __builtin_memcpy(&bits, &val[q], sizeof(bits));
        if constexpr (NTS) __builtin_nontemporal_store(bits, zt + (size_t)M * ZT * q);
        else zt[(size_t)M * ZT * q] = bits;
      }
    } else {
      unsigned* zt = reinterpret_cast<unsigned*>(Z) + zoff;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        if constexpr (NTS) __builtin_nontemporal_store(pack_half2<ZST>(val[q], zscale), zt + (size_t)M * ZT * q);
        else zt[(size_t)M * ZT * q] = pack_half2<ZST>(val[q], zscale);
      }
    }
  };
  for (long long u = u_lo + l; u < u_hi; u += nbx) {
    const int b = (int)(u / npairs), pair = (int)(u - (long long)b * npairs);
    const int kx = pair * KC + c;
    const bool valid = kx < Wh;
    // re-derived every unit from a register the compiler cannot see through, so the address and
    // twiddle computations are not hoisted out of the loop (they would spill)
    int j = tid / KC;
    asm volatile("" : "+v"(j));
    const int k2 = j % 16;
    cf* const colbuf = lds + c * CS;
    cf* const p_s1 = colbuf + 17 * j;                         // stage-1 stores: + q
    cf* const p_rd = colbuf + j + (j >> 4);                   // strided reads:  + 272 q
    cf* const p_s2 = colbuf + (j / 16) * 272 + k2;            // stage-2 stores: + 17 q
    const cf* ic = inpT + (size_t)(valid ? kx : 0) * H + j;   // input spectrum: + 256 q
    const size_t rb = (size_t)b * T * Wh + kx;
    // forward stage 1 (Ns = 1) has the single non-zero input x[j] (T <= 256), so stage 2 butterfly j
    // takes x[j / 16 + 16 q] straight from the rows (as k_mid4096 with DIRECT)
    cf v[16], iv[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int r = j / 16 + 16 * q;
      v[q] = (valid && r < T) ? row_elem<ZST>(rows, rb + (size_t)r * Wh, rinv) : cmake(0.f, 0.f);
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) iv[q] = valid ? ic[M * q] : cmake(0.f, 0.f);
    if constexpr (ROT) {
      asm volatile("" ::: "memory");  // the loads above are issued before the stores below
      if (have_prev) store_z(zp, zoff_prev);
    }
    // forward stage 2 (Ns = 16)
#pragma unroll
    for (int q = 1; q < 16; ++q) v[q] = cmul(v[q], twiddle<-1>(tw2q, 16 * q + k2));
    dft16<-1>(v);
#pragma unroll
    for (int q = 0; q < 16; ++q) p_s2[17 * q] = v[q];
    lds_barrier();
    // forward stage 3 (Ns = 256): X[j + 256 q]
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = p_rd[272 * q];
#pragma unroll
    for (int q = 1; q < 16; ++q)
      v[q] = cmul(v[q], twiddle<-1>(tw256, (j / 16) * q), twiddle<-1>(tw3q, 16 * q + k2));
    dft16<-1>(v);
    // ZeroCentralPixel and conj(reference) * input
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const cf xx = (kx == 0 && j == 0 && q == 0) ? cmake(0.f, 0.f) : v[q];
      v[q] = cconjmul(xx, iv[q]);
    }
    lds_barrier();
    // inverse stage 1
    dft16<+1>(v);
#pragma unroll
    for (int q = 0; q < 16; ++q) p_s1[q] = v[q];
    lds_barrier();
    // inverse stage 2
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = p_rd[272 * q];
    lds_barrier();
#pragma unroll
    for (int q = 1; q < 16; ++q) v[q] = cmul(v[q], twiddle<+1>(tw2q, 16 * q + k2));
    dft16<+1>(v);
#pragma unroll
    for (int q = 0; q < 16; ++q) p_s2[17 * q] = v[q];
    lds_barrier();
    // inverse stage 3: z[j + 256 q]
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = p_rd[272 * q];
#pragma unroll
    for (int q = 1; q < 16; ++q)
      v[q] = cmul(v[q], twiddle<+1>(tw256, (j / 16) * q), twiddle<+1>(tw3q, 16 * q + k2));
    dft16<+1>(v);
    const size_t zoff = b * z_bs + (size_t)(kx / ZT) * H * ZT + (size_t)j * ZT + kx % ZT;
    if constexpr (ROT) {
#pragma unroll
      for (int q = 0; q < 16; ++q) zp[q] = v[q];
      zoff_prev = zoff;
      have_prev = true;
    } else {
      store_z(v, zoff);
    }
    lds_barrier();  // LDS reads of this unit complete before the next one writes
  }
  if (ROT && have_prev) store_z(zp, zoff_prev);
}

// ---------------------------------------------------------------- row pass + epilogue
struct EpiParams {
  float* mip;
  int* idx;
  double* sum;
  double* sumsq;
  unsigned long long* hist;
  double hmin, hstep_inv, hstep;
  int nbins;
  int g0;      // global orientation index of batch entry 0
  const int* plane_g0;  // when set, *plane_g0 is added to g0 (a replayed graph reads its plane here)
  int ztile;   // column-tile width of Z (power of two)
  int ntiles;  // column tiles per orientation
  float zinv;  // inverse of the scale of 16-bit Z storage (1 for FP32)
};

// Histogram bins of correlation values as match_template.cpp:1124 computes them,
// int(double(v - histogram_min_scaled) / histogram_step_scaled) with the min and step in double.
// Bins outside 0..nbins-1 go to the spill bin nbins (NBIN_SPILL slots are reserved for it), which is
// never flushed: the reference's guard admits bin nbins, one past its array; here it is dropped.
// The quotient is formed with the reciprocal of the step, which can differ from the correctly
// rounded division by an ulp; bin_near_edge flags a quotient within 1e-9 of an integer (where that
// ulp could cross a bin edge), and the caller then moves the counts of values whose bin_exact
// differs, so every bin equals the reference expression's. A quotient in (-1, 0) truncates to bin 0, as in the reference.
__device__ __forceinline__ double bin_quotient(float v, const EpiParams& ep) {
  return ((double)v - ep.hmin) * ep.hstep_inv;
}
__device__ __forceinline__ bool bin_near_edge(double d) { return fabs(d - rint(d)) < 1e-9; }
__device__ __forceinline__ int bin_of(double d, const EpiParams& ep) {
  const int k = (int)d;  // truncation, as int() in the reference
  return (unsigned)k < (unsigned)ep.nbins ? k : ep.nbins;
}
__device__ __forceinline__ int bin_exact(float v, const EpiParams& ep) {
  return bin_of(((double)v - ep.hmin) / ep.hstep, ep);
}

// Threads of a k_last block: QR orientations x (W/2) / R1 threads, R1 the first radix of the
// W/2-point plan (radix 8 wherever W/2 has three factors of two).
template <int W, int QR>
__host__ __device__ constexpr int last_threads() { return QR * ((W / 2) / plan_radix(W / 2, 0, 3)); }

// Row pass + epilogue. One block per image row y; QR groups of NT = (W/2)/R1 threads transform QR
// orientations of the batch at a time: C2R of row y of Z into W reals (an h = W/2 point complex
// inverse FFT after the even/odd pre-twist), then every thread updates its 2*PAIRS pixels for
// those orientations in batch order: strict-greater MIP/argmax, histogram bin
// int(double(v - min) / step) over 0..nbins-1, FP64 sum += v and FP64 sum-of-squares += (v*v in float)
// (match_template.cpp:1100-1149). The per-pixel state stays in registers for the whole batch.
// Rows are assigned so that consecutive rows run on the same XCD and share Z's cache lines in L2.
// Twiddles live in LDS; the block is sized so that two blocks share a CU and overlap each other's
// Z gathers.
template <int W, int QR, int ZST = ZS_F32>
__global__ __launch_bounds__((last_threads<W, QR>())) __attribute__((amdgpu_waves_per_eu(4)))
void k_last(const cf* __restrict__ Z, int H, int nb, EpiParams ep, const cf* __restrict__ tw_n,
            const cf* __restrict__ tw_h) {
  constexpr int h = W / 2;
  constexpr int R1 = plan_radix(h, 0, 3);
  constexpr int NT = h / R1;  // one first-stage butterfly per thread
  constexpr int BT = QR * NT;
  constexpr int PAIRS = h / BT;
  static_assert(BT == last_threads<W, QR>() && h % BT == 0 && PAIRS >= 1, "QR must divide the first radix");
  constexpr int HC = 4;  // interleaved histogram copies (bin * HC + lane % HC) spread same-bin atomics
  __shared__ cf buf[QR][lds_len(h + 1)];
  __shared__ cf twh[h];
  __shared__ cf twn[h / 2 + 1];
  __shared__ unsigned int hist[513 * HC];  // bins 0..511 and the spill bin
  const int tid0 = threadIdx.x;
  const int bid = blockIdx.x;
  const int y = xcd_row(bid, H);
  const int gbase = ep.g0 + (ep.plane_g0 ? *ep.plane_g0 : 0);
  for (int i = tid0; i < 513 * HC; i += BT) hist[i] = 0u;
  for (int i = tid0; i < h; i += BT) twh[i] = tw_h[i];
  for (int i = tid0; i <= h / 2; i += BT) twn[i] = tw_n[i];
  float mip[2 * PAIRS];
  int idx[2 * PAIRS];
  double s1[2 * PAIRS], s2[2 * PAIRS];
  const size_t rowoff = (size_t)y * W;
#pragma unroll
  for (int u = 0; u < PAIRS; ++u) {
    const int m = tid0 + u * BT;
    const float2 a = *reinterpret_cast<const float2*>(ep.mip + rowoff + 2 * m);
    const int2 ia = *reinterpret_cast<const int2*>(ep.idx + rowoff + 2 * m);
    const double2 sa = *reinterpret_cast<const double2*>(ep.sum + rowoff + 2 * m);
    const double2 qa = *reinterpret_cast<const double2*>(ep.sumsq + rowoff + 2 * m);
    mip[2 * u] = a.x; mip[2 * u + 1] = a.y;
    idx[2 * u] = ia.x; idx[2 * u + 1] = ia.y;
    s1[2 * u] = sa.x; s1[2 * u + 1] = sa.y;
    s2[2 * u] = qa.x; s2[2 * u + 1] = qa.y;
  }
  const int sh = __ffs(ep.ztile) - 1;
  const size_t z_bs = (size_t)ep.ntiles * H * ep.ztile;
  for (int b0 = 0; b0 < nb; b0 += QR) {
    // re-derived every iteration from a register the compiler cannot see through, so address
    // arithmetic is not hoisted out of the loop into registers (it would spill)
    int tid = tid0;
    asm volatile("" : "+v"(tid));
    const int grp = tid / NT, t = tid % NT;
    const int nbp = min(QR, nb - b0);
    const bool active = grp < nbp;
    cf* const bb = buf[grp];
    // gather row y of orientation b0 + grp (elements 0..h) from the column tiles
    if (active) {
      const size_t zo = (size_t)(b0 + grp) * z_bs + (size_t)y * ep.ztile;
      if constexpr (ZST == ZS_F32) {
        const cf* zb = Z + zo;
        for (int e = t; e <= h; e += NT) bb[lpad(e)] = zb[(size_t)(e >> sh) * H * ep.ztile + (e & (ep.ztile - 1))];
      } else {
        const unsigned* zb = reinterpret_cast<const unsigned*>(Z) + zo;
        for (int e = t; e <= h; e += NT)
          bb[lpad(e)] = unpack_half2<ZST>(zb[(size_t)(e >> sh) * H * ep.ztile + (e & (ep.ztile - 1))], ep.zinv);
      }
    }
    __syncthreads();
    // first radix-8 stage (Ns = 1) with the real-to-complex pre-twist folded in:
    // Z_k = (X_k + conj X_{h-k}) + i (X_k - conj X_{h-k}) conj(W_W^k), with W_W^k = -conj(W_W^{h-k})
    // for k > h/2, and Z_0 = (Re X_0 + Re X_h) + i (Re X_0 - Re X_h)
    cf v[R1];
    if (active) {
#pragma unroll
      for (int q = 0; q < R1; ++q) {
        const int k = t + NT * q;
        const cf Xk = bb[lpad(k)];
        const cf Xm = bb[lpad(h - k)];
        if (k == 0) {
          v[q] = cmake(Xk.x + Xm.x, Xk.x - Xm.x);
        } else {
          const cf Xc = cconj(Xm);
          const cf E = cadd(Xk, Xc);
          const cf D = csub(Xk, Xc);
          cf w;
          if (k <= h / 2) {
            w = twn[k];
          } else {
            const cf w2 = twn[h - k];
            w = cmake(-w2.x, w2.y);
          }
          const cf O = cmul(D, cconj(w));
          v[q] = cmake(E.x - O.y, E.y + O.x);
        }
      }
    }
    __syncthreads();
    if (active) {
      dftR<R1, +1>(v);
#pragma unroll
      for (int q = 0; q < R1; ++q) bb[lpad(t * R1 + q)] = v[q];
    }
    __syncthreads();
    // remaining stages (the groups past nbp run them on stale data; their result is unused)
    fft_inplace<h, +1, NT, 3, 1>(bb, twh, t);
#pragma unroll 1
    for (int q = 0; q < nbp; ++q) {
      const int g = gbase + b0 + q;
      bool near = false;
#pragma unroll
      for (int u = 0; u < PAIRS; ++u) {
        const cf z = buf[q][lpad(tid + u * BT)];
        const float vv[2] = {z.x, z.y};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const float vx = vv[c];
          const int p = 2 * u + c;
          if (vx > mip[p]) { mip[p] = vx; idx[p] = g; }
          const double d = bin_quotient(vx, ep);
          atomicAdd(&hist[bin_of(d, ep) * HC + tid % HC], 1u);
          near |= bin_near_edge(d);
          s1[p] += (double)vx;
          const float sq = vx * vx;
          s2[p] += (double)sq;
        }
      }
      if (__builtin_expect(near, 0)) {
        // rare: move each value whose exact quotient falls in another bin
#pragma unroll 1
        for (int p = 0; p < 2 * PAIRS; ++p) {
          const cf z = buf[q][lpad(tid + (p >> 1) * BT)];
          const float vx = (p & 1) ? z.y : z.x;
          const int a = bin_of(bin_quotient(vx, ep), ep), e = bin_exact(vx, ep);
          if (a != e) {
            atomicSub(&hist[a * HC + tid % HC], 1u);
            atomicAdd(&hist[e * HC + tid % HC], 1u);
          }
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < PAIRS; ++u) {
    const int m = tid0 + u * BT;
    *reinterpret_cast<float2*>(ep.mip + rowoff + 2 * m) = make_float2(mip[2 * u], mip[2 * u + 1]);
    *reinterpret_cast<int2*>(ep.idx + rowoff + 2 * m) = make_int2(idx[2 * u], idx[2 * u + 1]);
    *reinterpret_cast<double2*>(ep.sum + rowoff + 2 * m) = make_double2(s1[2 * u], s1[2 * u + 1]);
    *reinterpret_cast<double2*>(ep.sumsq + rowoff + 2 * m) = make_double2(s2[2 * u], s2[2 * u + 1]);
  }
  for (int i = tid0; i < ep.nbins; i += BT) {
    unsigned long long n = 0;
#pragma unroll
    for (int c = 0; c < HC; ++c) n += hist[i * HC + c];
    if (n) atomicAdd(&ep.hist[i], n);
  }
}

typedef __attribute__((address_space(3))) void* lds_vptr;
typedef const __attribute__((address_space(1))) void* global_vptr;

// k_last with the Z gather double-buffered through direct global-to-LDS loads
// (global_load_lds_dwordx4, 16 bytes = two FP32 or four 16-bit adjacent elements of a tile row per
// lane; 16-bit elements are widened and unscaled when the first stage reads them): while the block transforms
// and accumulates orientations b0..b0+QR-1, row y of the next QR orientations streams into a second
// LDS buffer. Pixel ownership is that of k_last; the pre-twist and FFT twiddles come from factored
// tables so that two blocks still fit in one CU's LDS, except those of the stages whose angles are
// multiples of 2 pi / 64, read from a direct 64-entry table, and in the later stages each
// butterfly looks up its first twiddle and forms the others as its powers (the values then agree
// with k_last's to FP32 rounding, not bit for bit). Each group of NT = W/16 threads
// must be whole wavefronts (W >= 1024), because a direct load writes one wavefront's 64 lanes to
// consecutive LDS addresses.
template <int W, int QR, int ZST = ZS_F32>
__global__ __launch_bounds__(QR * (W / 16)) __attribute__((amdgpu_waves_per_eu(4)))
void k_last_dma(const cf* __restrict__ Z, int H, int nb, EpiParams ep, const cf* __restrict__ tw_n,
                const cf* __restrict__ tw_h) {
  constexpr int h = W / 2;
  constexpr int R1 = plan_radix(h, 0, 3);
  constexpr int NT = h / R1;
  constexpr int BT = QR * NT;
  constexpr int PAIRS = h / BT;
  constexpr int EPL = ZST == ZS_F32 ? 2 : 4;  // Z elements per 16-byte load (within one tile row)
  constexpr int ESZ = 16 / EPL;               // bytes per stored element
  constexpr int NL = h / EPL + 1;             // 16-byte loads per row: elements 0..h plus padding
  constexpr int NR = (NL + NT - 1) / NT;      // load rounds per thread
  static_assert(R1 == 8 && h % BT == 0 && NT % 64 == 0, "groups of whole wavefronts, one butterfly each");
  constexpr int HC = 4;
  __shared__ __align__(16) unsigned char gbuf[QR][16 * NL];  // gathered rows in element order
  __shared__ cf buf[QR][lds_len(h)];       // FFT work buffers (padded index)
  __shared__ cf fft_lo[64], fft_hi[h / 64], fft_d64[64];
  __shared__ cf pre_lo[64], pre_hi[h / 128 + 1];
  __shared__ unsigned int hist[513 * HC];  // bins 0..511 and the spill bin
  const int tid0 = threadIdx.x;
  const int bid = blockIdx.x;
  const int y = xcd_row(bid, H);
  const int gbase = ep.g0 + (ep.plane_g0 ? *ep.plane_g0 : 0);
  const int sh = __ffs(ep.ztile) - 1;
  const size_t z_bs = (size_t)ep.ntiles * H * ep.ztile;
  const size_t tile_stride = (size_t)H * ep.ztile;
  const unsigned char* const zrow = reinterpret_cast<const unsigned char*>(Z) + (size_t)y * ep.ztile * ESZ;
  // start the loads of row y of orientations b..b+QR-1 (group grp takes orientation b + grp); one
  // load takes EPL consecutive elements of a tile row (the tile width is at least EPL)
  auto gather = [&](int b, int tid) {
    const int grp = tid / NT, t = tid % NT;
    if (b + grp < nb) {
      const unsigned char* zb = zrow + (size_t)(b + grp) * z_bs * ESZ;
      unsigned char* const dst = gbuf[grp] + 16 * (t & ~63);
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const int L = r * NT + t;
        if (L < NL) {
          const int e = EPL * L;
          const unsigned char* src = zb + ((size_t)(e >> sh) * tile_stride + (e & (ep.ztile - 1))) * ESZ;
          __builtin_amdgcn_global_load_lds((global_vptr)src, (lds_vptr)(dst + 16 * r * NT), 16, 0, 0);
        }
      }
    }
  };
  // element k of a gathered row
  auto zload = [&](const unsigned char* gb, int k) -> cf {
    if constexpr (ZST == ZS_F32) return reinterpret_cast<const cf*>(gb)[k];
    else return unpack_half2<ZST>(reinterpret_cast<const unsigned*>(gb)[k], ep.zinv);
  };
  gather(0, tid0);
  for (int i = tid0; i < 513 * HC; i += BT) hist[i] = 0u;
  for (int i = tid0; i < 64; i += BT) {
    fft_lo[i] = tw_h[i];
    pre_lo[i] = tw_n[i];
  }
  for (int i = tid0; i < h / 64; i += BT) fft_hi[i] = tw_h[64 * i];
  for (int i = tid0; i < 64; i += BT) fft_d64[i] = tw_h[(h / 64) * i];
  for (int i = tid0; i <= h / 128; i += BT) pre_hi[i] = tw_n[64 * i];
  float mip[2 * PAIRS];
  int idx[2 * PAIRS];
  double s1[2 * PAIRS], s2[2 * PAIRS];
  const size_t rowoff = (size_t)y * W;
#pragma unroll
  for (int u = 0; u < PAIRS; ++u) {
    const int m = tid0 + u * BT;
    const float2 a = *reinterpret_cast<const float2*>(ep.mip + rowoff + 2 * m);
    const int2 ia = *reinterpret_cast<const int2*>(ep.idx + rowoff + 2 * m);
    const double2 sa = *reinterpret_cast<const double2*>(ep.sum + rowoff + 2 * m);
    const double2 qa = *reinterpret_cast<const double2*>(ep.sumsq + rowoff + 2 * m);
    mip[2 * u] = a.x; mip[2 * u + 1] = a.y;
    idx[2 * u] = ia.x; idx[2 * u + 1] = ia.y;
    s1[2 * u] = sa.x; s1[2 * u + 1] = sa.y;
    s2[2 * u] = qa.x; s2[2 * u + 1] = qa.y;
  }
  const TwHybrid ftw{fft_d64, TwSplit{fft_lo, fft_hi}};
  const TwSplit ptw{pre_lo, pre_hi};
  for (int b0 = 0; b0 < nb; b0 += QR) {
    // re-derived every iteration from a register the compiler cannot see through, so address
    // arithmetic is not hoisted out of the loop into registers (it would spill)
    int tid = tid0;
    asm volatile("" : "+v"(tid));
    const int grp = tid / NT, t = tid % NT;
    const int nbp = min(QR, nb - b0);
    const bool active = grp < nbp;
    // this wave's direct loads have landed (vmcnt 0); the barrier makes every wave's visible
    lds_dma_barrier();
    // first radix-8 stage (Ns = 1) with the real-to-complex pre-twist folded in:
    // Z_k = (X_k + conj X_{h-k}) + i (X_k - conj X_{h-k}) conj(W_W^k), with W_W^k = -conj(W_W^{h-k})
    // for k > h/2, and Z_0 = (Re X_0 + Re X_h) + i (Re X_0 - Re X_h)
    cf v[R1];
    if (active) {
      const unsigned char* gb = gbuf[grp];
      // W_W^k for k = t + NT q: the table's W_W^t times W_W^(NT q) = W_16^q (NT = W / 16), a constant
      static_assert(R1 == 8 && NT * 16 == W, "k = t + (W / 16) q");
      constexpr float c16[8] = {1.0f, 0.92387953251128675613f, 0.70710678118654752440f, 0.38268343236508977173f,
                                0.0f, -0.38268343236508977173f, -0.70710678118654752440f, -0.92387953251128675613f};
      constexpr float s16[8] = {0.0f, 0.38268343236508977173f, 0.70710678118654752440f, 0.92387953251128675613f,
                                1.0f, 0.92387953251128675613f, 0.70710678118654752440f, 0.38268343236508977173f};
      const cf wt = twiddle<-1>(ptw, t);
#pragma unroll
      for (int q = 0; q < R1; ++q) {
        const int k = t + NT * q;
        const cf Xk = zload(gb, k);
        const cf Xm = zload(gb, h - k);
        if (k == 0) {
          v[q] = cmake(Xk.x + Xm.x, Xk.x - Xm.x);
        } else {
          const cf Xc = cconj(Xm);
          const cf E = cadd(Xk, Xc);
          const cf D = csub(Xk, Xc);
          const cf w = q == 0 ? wt : cmul(wt, cmake(c16[q], -s16[q]));
          const cf O = cmul(D, cconj(w));
          v[q] = cmake(E.x - O.y, E.y + O.x);
        }
      }
    }
    lds_barrier();  // every read of gbuf is done: the next orientations may overwrite it
    gather(b0 + QR, tid);
    if (active) {
      dftR<R1, +1>(v);
#pragma unroll
      for (int q = 0; q < R1; ++q) buf[grp][lpad(t * R1 + q)] = v[q];
    }
    lds_barrier();
    // remaining stages (groups past nbp run them on stale data; their result is unused)
    fft_inplace<h, +1, NT, 3, 1>(buf[grp], ftw, t);
#pragma unroll 1
    for (int q = 0; q < nbp; ++q) {
      const int g = gbase + b0 + q;
      bool near = false;
#pragma unroll
      for (int u = 0; u < PAIRS; ++u) {
        const cf z = buf[q][lpad(tid + u * BT)];
        const float vv[2] = {z.x, z.y};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const float vx = vv[c];
          const int p = 2 * u + c;
          if (vx > mip[p]) { mip[p] = vx; idx[p] = g; }
          const double d = bin_quotient(vx, ep);
          lds_add(&hist[bin_of(d, ep) * HC + tid % HC], 1u);
          near |= bin_near_edge(d);
          s1[p] += (double)vx;
          const float sq = vx * vx;
          s2[p] += (double)sq;
        }
      }
      if (__builtin_expect(near, 0)) {
        // rare: move each value whose exact quotient falls in another bin
#pragma unroll 1
        for (int p = 0; p < 2 * PAIRS; ++p) {
          const cf z = buf[q][lpad(tid + (p >> 1) * BT)];
          const float vx = (p & 1) ? z.y : z.x;
          const int a = bin_of(bin_quotient(vx, ep), ep), e = bin_exact(vx, ep);
          if (a != e) {
            lds_add(&hist[a * HC + tid % HC], 0xFFFFFFFFu);  // minus one, modulo 2^32
            lds_add(&hist[e * HC + tid % HC], 1u);
          }
        }
      }
    }
    // no barrier here: the next iteration's first barrier orders these reads of buf before the
    // next first stage rewrites it
  }
  lds_barrier();
#pragma unroll
  for (int u = 0; u < PAIRS; ++u) {
    const int m = tid0 + u * BT;
    *reinterpret_cast<float2*>(ep.mip + rowoff + 2 * m) = make_float2(mip[2 * u], mip[2 * u + 1]);
    *reinterpret_cast<int2*>(ep.idx + rowoff + 2 * m) = make_int2(idx[2 * u], idx[2 * u + 1]);
    *reinterpret_cast<double2*>(ep.sum + rowoff + 2 * m) = make_double2(s1[2 * u], s1[2 * u + 1]);
    *reinterpret_cast<double2*>(ep.sumsq + rowoff + 2 * m) = make_double2(s2[2 * u], s2[2 * u + 1]);
  }
  for (int i = tid0; i < ep.nbins; i += BT) {
    unsigned long long n = 0;
#pragma unroll
    for (int c = 0; c < HC; ++c) n += hist[i * HC + c];
    if (n) atomicAdd(&ep.hist[i], n);
  }
}

// ---------------------------------------------------------------- state and finalisation
static __global__ void k_reset_state(float* mip, int* idx, double* sum, double* sumsq, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  mip[i] = -FLT_MAX;
  idx[i] = -1;
  sum[i] = 0.0;
  sumsq[i] = 0.0;
}

static __global__ void k_zero_u64(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// AggregatedTemplateResult::AddResult (match_template.cpp:2015-2103) for two partial searches of
// one image held on one device: the larger MIP wins with its orientation; with global orientation
// indices an equal MIP keeps the lower index, the orientation the sequential loop meets first, so
// merging position ranges reproduces one search over their union. Sums add.
static __global__ void k_merge(float* __restrict__ mip, int* __restrict__ idx, double* __restrict__ sum,
                               double* __restrict__ sumsq, const float* __restrict__ smip,
                               const int* __restrict__ sidx, const double* __restrict__ ssum,
                               const double* __restrict__ ssumsq, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float a = mip[i], b = smip[i];
  const int ia = idx[i], ib = sidx[i];
  if (ib >= 0 && (ia < 0 || b > a || (b == a && ib < ia))) {
    mip[i] = b;
    idx[i] = ib;
  }
  sum[i] += ssum[i];
  sumsq[i] += ssumsq[i];
}

static __global__ void k_add_u64(unsigned long long* __restrict__ a, const unsigned long long* __restrict__ b, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] += b[i];
}

// One pixel of the finalisation: match_template.cpp:1221-1243, 1300-1310 with the reference's
// float/double mix (scaled MIP, z-score, average and standard deviation).
struct FinalPixel {
  float m, z, avg, sd;
};
__device__ inline FinalPixel finalize_pixel(float mip, double sum, double sumsq, float n_ccf, float sqrt_px) {
  double s = sum / n_ccf;
  double q = sumsq / n_ccf - (double)powf((float)s, 2);
  if (q > 0.0f) q = (double)(sqrtf((float)q) * sqrt_px);
  else q = 0.0f;
  s *= sqrt_px;
  const float m = mip * sqrt_px;
  float zz = (float)((double)m - s);
  if (q > 0.0f) zz = (float)((double)zz / q);
  else zz = 0.0f;
  return {m, zz, (float)s, (float)q};
}

static __global__ void k_finalize(const float* __restrict__ mip, const double* __restrict__ sum,
                           const double* __restrict__ sumsq, size_t n, float n_ccf, float sqrt_px,
                           float* __restrict__ mip_scaled, float* __restrict__ z, float* __restrict__ avg,
                           float* __restrict__ sd) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const FinalPixel f = finalize_pixel(mip[i], sum[i], sumsq[i], n_ccf, sqrt_px);
  mip_scaled[i] = f.m;
  z[i] = f.z;
  avg[i] = f.avg;
  sd[i] = f.sd;
}

}  // namespace tmk
