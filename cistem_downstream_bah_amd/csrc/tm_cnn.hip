// Proposal network (include/tm_cnn.h): 3x3 convolutions with ReLU and a 1x1 head, forward,
// backward and Adam on the device. No counterpart in the reference.
//
// Products run on bf16 MFMA tiles with FP32 accumulation, channels-last FP32 activations and
// gradients read straight into the product (no im2col matrix, no transposes) wherever the channel
// counts allow:
//   forward        Y[pixel][cout] = relu(col(pixel, k) . W[cout][k] + b)   k_conv_fwd (implicit col)
//   weight grad    dW[cout][k] = sum over pixels of dZ[pixel][cout] col(pixel, k)   k_conv_wgrad
//   data grad      dX[pixel][cin] = sum over the taps reaching it of dZ W          k_conv_dgrad
// the 1x1 head with the same kernels (one tap), the first layer's weight gradient with its nine taps
// as the product's n. Weights go to the products packed in MFMA operand order (k_pack_b). Layers
// whose channel counts the implicit kernels do not take use explicit products (k_gemm_nt / k_gemm_nt4
// over im2col and transposed operands, col2im). Every split reduction adds its partials in a fixed
// order, so results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "../../include/tm_cnn.h"

namespace {

thread_local std::string g_cerr;

struct CnnError {
  int code;
  std::string msg;
};

#define CNN_HIP(call)                                                                                  \
  do {                                                                                                 \
    hipError_t e_ = (call);                                                                            \
    if (e_ != hipSuccess) throw CnnError{TM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
  } while (0)
#define CNN_REQUIRE(cond, code, msg)                                                                   \
  do {                                                                                                 \
    if (!(cond)) throw CnnError{code, msg};                                                            \
  } while (0)

template <class F>
int cguard(F&& f) {
  try {
    f();
    return TM_OK;
  } catch (const CnnError& e) {
    g_cerr = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    g_cerr = e.what();
    return TM_ERR_INVALID;
  }
}

void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw CnnError{TM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e)};
}

typedef unsigned short bf16_t;
typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// float -> bf16, round to nearest even (NaN kept quiet)
__device__ inline bf16_t f2bf(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (bf16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
__device__ inline float bf2f(bf16_t b) { return __uint_as_float((unsigned)b << 16); }
// *p, or `zero` when !ok; p must be a valid address either way. Loading first and selecting after keeps
// the compiler from wrapping each masked load in a branch of its own.
template <class T>
__device__ inline T load_or(const T* p, bool ok, T zero) {
  const T v = *p;
  return ok ? v : zero;
}
__device__ inline float load_f(const float* p) { return *p; }
__device__ inline float load_f(const bf16_t* p) { return bf2f(*p); }

constexpr int round16(int x) { return (x + 15) / 16 * 16; }
inline long long round16ll(long long x) { return (x + 15) / 16 * 16; }
inline unsigned grid1(long long n, int bs = 256, long long cap = 1 << 20) {
  long long g = (n + bs - 1) / bs;
  return (unsigned)std::max(1LL, std::min(g, cap));
}

// ------------------------------------------------------------------ kernels
// C (+ z * split_stride) = A[M][lda] . B[N][ldb]^T over k in [z * kchunk, min(K, (z + 1) * kchunk)),
// K and kchunk multiples of 16, lda and ldb multiples of 4. A block of four waves covers
// (128 / WN) rows x (32 WN) columns; wave w owns rows 32 (w / WN) .. and columns 32 (w % WN) .. as
// 2 x 2 MFMA tiles (WN = 2: 64 x 64; WN = 1: 128 x 32 for N <= 32; WN = 4: 32 x 128 for M <= 32).
// MFMA 16x16x16 bf16 lane layout: A[m = lane % 16][k = 4 (lane / 16) + i],
// B[n = lane % 16][k = 4 (lane / 16) + i], D[m = 4 (lane / 16) + i][n = lane % 16].
template <bool BIAS, bool RELU, int WN>
__global__ __launch_bounds__(256) void k_gemm_nt(const bf16_t* __restrict__ A, int lda, const bf16_t* __restrict__ B,
                                                 int ldb, float* __restrict__ C, long long ldc, long long M, int N,
                                                 int K, int kchunk, long long split_stride,
                                                 const float* __restrict__ bias) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long m0 = (long long)blockIdx.x * (128 / WN) + (wave / WN) * 32;
  const int n0 = blockIdx.y * (32 * WN) + (wave % WN) * 32;
  const int r = lane & 15, kq = (lane >> 4) * 4;
  const int k_lo = blockIdx.z * kchunk, k_hi = min(K, k_lo + kchunk);
  const bf16_t* ap[2];
  const bf16_t* bp[2];
  bool aok[2], bok[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const long long m = m0 + i * 16 + r;
    aok[i] = m < M;
    ap[i] = A + (aok[i] ? m : 0) * (long long)lda + kq;
    const int n = n0 + i * 16 + r;
    bok[i] = n < N;
    bp[i] = B + (size_t)(bok[i] ? n : 0) * ldb + kq;
  }
  float4_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
  const short4_t zero = {0, 0, 0, 0};
  // KU k steps per iteration: every load of the group is issued before its MFMAs
  constexpr int KU = 4;
  int k = k_lo;
  for (; k + 16 * KU <= k_hi; k += 16 * KU) {
    short4_t a[KU][2], b[KU][2];
#pragma unroll
    for (int u = 0; u < KU; ++u)
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        a[u][i] = load_or(reinterpret_cast<const short4_t*>(ap[i] + k + 16 * u), aok[i], zero);
        b[u][i] = load_or(reinterpret_cast<const short4_t*>(bp[i] + k + 16 * u), bok[i], zero);
      }
#pragma unroll
    for (int u = 0; u < KU; ++u)
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[u][i], b[u][j], acc[i][j], 0, 0, 0);
  }
  for (; k < k_hi; k += 16) {
    short4_t a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      a[i] = load_or(reinterpret_cast<const short4_t*>(ap[i] + k), aok[i], zero);
      b[i] = load_or(reinterpret_cast<const short4_t*>(bp[i] + k), bok[i], zero);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
  }
  float* Cz = C + (long long)blockIdx.z * split_stride;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const long long m = m0 + i * 16 + kq + e;
        const int n = n0 + j * 16 + r;
        if (m < M && n < N) {
          float v = acc[i][j][e];
          if (BIAS) v += bias[n];
          if (RELU) v = fmaxf(v, 0.f);
          Cz[m * ldc + n] = v;
        }
      }
}

// The same product for M >= 128 and N >= 128 on a 128 x 128 block tile: each wave owns 64 x 64 as 4 x 4
// MFMA tiles (every loaded operand feeds four MFMAs) and steps k by 32 on the 16x16x32 bf16 MFMA, a
// lane's eight values of a row being one 16-byte load (A and B use the same lane-to-k assignment, so
// the product pairs them whatever order the hardware sums its 32 k in); a 16-wide k remainder takes one
// 16x16x16 step. K, kchunk multiples of 16; lda, ldb multiples of 8.
typedef short short8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
template <bool BIAS, bool RELU>
__global__ __launch_bounds__(256) void k_gemm_nt4(const bf16_t* __restrict__ A, int lda, const bf16_t* __restrict__ B,
                                                  int ldb, float* __restrict__ C, long long ldc, long long M, int N,
                                                  int K, int kchunk, long long split_stride,
                                                  const float* __restrict__ bias) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long m0 = (long long)blockIdx.x * 128 + (wave >> 1) * 64;
  const int n0 = blockIdx.y * 128 + (wave & 1) * 64;
  const int r = lane & 15, kg = lane >> 4;
  const int k_lo = blockIdx.z * kchunk, k_hi = min(K, k_lo + kchunk);
  const bf16_t* ap[4];
  const bf16_t* bp[4];
  bool aok[4], bok[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const long long m = m0 + i * 16 + r;
    aok[i] = m < M;
    ap[i] = A + (aok[i] ? m : 0) * (long long)lda;
    const int n = n0 + i * 16 + r;
    bok[i] = n < N;
    bp[i] = B + (size_t)(bok[i] ? n : 0) * ldb;
  }
  float4_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
  const short8_t zero8 = {0, 0, 0, 0, 0, 0, 0, 0};
  // two 32-wide k steps per iteration, their loads issued before their MFMAs
  int k = k_lo;
  for (; k + 64 <= k_hi; k += 64) {
    short8_t a[2][4], b[2][4];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int kk = k + 32 * u + 8 * kg;
        a[u][i] = load_or(reinterpret_cast<const short8_t*>(ap[i] + kk), aok[i], zero8);
        b[u][i] = load_or(reinterpret_cast<const short8_t*>(bp[i] + kk), bok[i], zero8);
      }
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a[u][i]),
                                                              __builtin_bit_cast(bf16x8_t, b[u][j]), acc[i][j], 0, 0, 0);
  }
  for (; k + 32 <= k_hi; k += 32) {
    short8_t a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a[i] = load_or(reinterpret_cast<const short8_t*>(ap[i] + k + 8 * kg), aok[i], zero8);
      b[i] = load_or(reinterpret_cast<const short8_t*>(bp[i] + k + 8 * kg), bok[i], zero8);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a[i]),
                                                            __builtin_bit_cast(bf16x8_t, b[j]), acc[i][j], 0, 0, 0);
  }
  if (k < k_hi) {  // 16 remaining
    const short4_t zero = {0, 0, 0, 0};
    short4_t a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a[i] = load_or(reinterpret_cast<const short4_t*>(ap[i] + k + 4 * kg), aok[i], zero);
      b[i] = load_or(reinterpret_cast<const short4_t*>(bp[i] + k + 4 * kg), bok[i], zero);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
  }
  float* Cz = C + (long long)blockIdx.z * split_stride;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = n0 + j * 16 + r;
      if (n >= N) continue;
      const float bn = BIAS ? bias[n] : 0.f;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const long long m = m0 + i * 16 + 4 * kg + e;
        if (m < M) {
          float v = acc[i][j][e] + bn;
          if (RELU) v = fmaxf(v, 0.f);
          Cz[m * ldc + n] = v;
        }
      }
    }
}

// Block x index dealt to the XCDs in contiguous ranges: block b runs on XCD b % 8, so when the x
// extent is a multiple of 8 the blocks of XCD q take pixel blocks q n / 8 .. (q + 1) n / 8 - 1 (all
// their n tiles too), and the nine taps' rereads of neighbouring pixels meet in that XCD's L2.
__device__ inline int xcd_block_x() {
  const int n = gridDim.x, b = blockIdx.x;
  return (n & 7) ? b : (b & 7) * (n >> 3) + (b >> 3);
}

// 3x3 convolution kernels as implicit products on MFMA 16x16x16 bf16 tiles. A wave owns 64 pixels (four
// 16-row tiles) x 16 NT channels, so each loaded operand feeds NT or four MFMAs; a block is four waves
// along the pixels. Every 16-wide k step lies in one tap, and a lane's four values of it are four
// consecutive channels of one pixel: one 16-byte load, rounded to bf16 in registers.
__device__ inline short4_t bf16x4_hw(float a, float b, float c, float d) {
  typedef __bf16 bf16x4v __attribute__((ext_vector_type(4)));
  const bf16x4v v = {(__bf16)a, (__bf16)b, (__bf16)c, (__bf16)d};
  return __builtin_bit_cast(short4_t, v);
}
__device__ inline short4_t bf16x4_hw(float4 v) { return bf16x4_hw(v.x, v.y, v.z, v.w); }

// Weights packed in MFMA operand order: Bp[j][s][lane][e] = bf16(B(n = 16 j + lane % 16,
// k = 16 s + 4 (lane / 16) + e)), zero for n >= N, for j < NBp (a multiple of 4 blocks, so that a
// wave's four 16-row tiles always lie inside) and s < K / 16: the 64 lanes of one operand load read
// 512 contiguous bytes. B(n, k) by MODE: PACK_ROWS W[n ldw + k] (k < Kw, else 0); PACK_DGRAD
// W[co 9 Cin + tap Cin + n] for k = tap Cout + co (the data gradient's B, the weights transposed per
// tap); PACK_COLS W[k ldw + n] (k < Kw, else 0; the head's data gradient).
enum { PACK_ROWS = 0, PACK_DGRAD = 1, PACK_COLS = 2 };
template <int MODE>
__global__ void k_pack_b(const float* __restrict__ W, int N, int K, int ldw, int Kw, int Cin, int Cout, int NBp,
                         bf16_t* __restrict__ Bp) {
  const int S = K / 16;
  const long long total = (long long)NBp * S * 256;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
    const int q = (int)(e & 3), lane = (int)((e >> 2) & 63);
    const long long js = e >> 8;
    const int s = (int)(js % S), j = (int)(js / S);
    const int n = 16 * j + (lane & 15), k = 16 * s + 4 * (lane >> 4) + q;
    float v = 0.f;
    if (n < N) {
      if (MODE == PACK_ROWS && k < Kw) v = W[(size_t)n * ldw + k];
      if (MODE == PACK_COLS && k < Kw) v = W[(size_t)k * ldw + n];
      if (MODE == PACK_DGRAD) {
        const int tap = k / Cout, co = k - tap * Cout;
        v = W[(size_t)co * 9 * Cin + tap * Cin + n];
      }
    }
    Bp[e] = f2bf(v);
  }
}

// The A operand of the convolutions: lane L loads four channels (4 (L % 4) ..) of pixel L / 4 of a
// 16-pixel tile, so four consecutive lanes read one pixel's 64 contiguous bytes, and ds_bpermute then
// hands MFMA lane (r, g) = r + 16 g the values of pixel r, channels 4 g ..: those of lane 4 r + g.
__device__ inline short4_t to_mfma_lanes(float4 v, int src_addr) {
  const short4_t h = bf16x4_hw(v);
  int2 w = __builtin_bit_cast(int2, h);
  w.x = __builtin_amdgcn_ds_bpermute(src_addr, w.x);
  w.y = __builtin_amdgcn_ds_bpermute(src_addr, w.y);
  return __builtin_bit_cast(short4_t, w);
}

// Forward, for Cin % 16 == 0: Y[m][n] = epi(sum_k col(m, k) B[n][k]) with
// col(m, k) = bf16(X[b][oy * stride + ky - 1][ox * stride + kx - 1][c]) for m = (b, oy, ox),
// k = (ky * 3 + kx) * Cin + c (0 outside the image), read from the FP32 activations instead of a
// materialised im2col matrix; TAPS = 1 takes the centre tap alone (a 1x1 layer, k = c: the head, and
// the head's data gradient with X = dL/dlogits). epi: EPI_BIAS_RELU relu(v + bias[n]), EPI_BIAS
// v + bias[n], EPI_MASK (A[m][n] > 0 ? v : 0, the backward of the ReLU that produced A). Bp: B packed
// by k_pack_b (K = TAPS Cin). Si^2 Cin < 2^31, M N < 2^31.
enum { EPI_BIAS_RELU = 0, EPI_BIAS = 1, EPI_MASK = 2 };
template <int NT, int TAPS, int EPI>
__global__ __launch_bounds__(256) void k_conv_fwd(const float* __restrict__ X, int Si, int Cin, int So, int stride,
                                                  const bf16_t* __restrict__ Bp, float* __restrict__ Y, int M, int N,
                                                  const float* __restrict__ bias, const float* __restrict__ A) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m0 = xcd_block_x() * 256 + wave * 64, n0 = blockIdx.y * 16 * NT;
  const int r = lane & 15, kq = (lane >> 4) * 4;
  const int pl = lane >> 2, cq = (lane & 3) * 4, src = (4 * r + (lane >> 4)) * 4;
  const int SoSo = So * So;
  int y0[4], x0[4];
  const float* xb[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + 16 * i + pl;
    const int mm = m < M ? m : 0;
    const int b = mm / SoSo, pix = mm - b * SoSo;
    const int oy = pix / So, ox = pix - oy * So;
    // rows past M get a start that no tap brings into the image
    y0[i] = m < M ? oy * stride - 1 : -4;
    x0[i] = ox * stride - 1;
    xb[i] = X + (size_t)b * Si * Si * Cin + cq;
  }
  const int S = TAPS * (Cin / 16);
  const bf16_t* bp = Bp + ((size_t)(n0 / 16) * S * 64 + lane) * 4;
  float4_t acc[4][NT];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
  const float4 zf4 = {0.f, 0.f, 0.f, 0.f};
  // k steps of 16 channels over (tap, c0), flattened: the loads of step s + 1 are in flight while the
  // MFMAs of step s run (two register buffers)
  int tap = TAPS == 9 ? 0 : 4, c0 = 0, ks = 0;
  const float* ap[4];
  bool aok[4];
  auto setup = [&]() {
    const int ky = tap / 3, kx = tap - 3 * (tap / 3);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int iy = y0[i] + ky, ix = x0[i] + kx;
      aok[i] = iy >= 0 && iy < Si && ix >= 0 && ix < Si;
      // a pixel clamped into the image: a valid address without a branch, its value then dropped
      ap[i] = xb[i] + (min(max(iy, 0), Si - 1) * Si + min(max(ix, 0), Si - 1)) * Cin;
    }
  };
  auto load = [&](float4 (&ar)[4], short4_t (&br)[NT]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) ar[i] = load_or(reinterpret_cast<const float4*>(ap[i] + c0), aok[i], zf4);
#pragma unroll
    for (int j = 0; j < NT; ++j) br[j] = *reinterpret_cast<const short4_t*>(bp + ((size_t)j * S + ks) * 256);
    ++ks;
    c0 += 16;
    if (c0 == Cin) {  // next tap (past the last one the addresses are computed and never used)
      c0 = 0;
      ++tap;
      setup();
    }
  };
  auto mma = [&](const float4 (&ar)[4], const short4_t (&br)[NT]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const short4_t a = to_mfma_lanes(ar[i], src);
#pragma unroll
      for (int j = 0; j < NT; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, br[j], acc[i][j], 0, 0, 0);
    }
  };
  setup();
  float4 a0[4], a1[4];
  short4_t b0[NT], b1[NT];
  load(a0, b0);
  int s = 1;  // the next step to load; a0 holds step s - 1
  for (; s + 1 < S; s += 2) {
    load(a1, b1);
    mma(a0, b0);
    load(a0, b0);
    mma(a1, b1);
  }
  if (s < S) {
    load(a1, b1);
    mma(a0, b0);
    mma(a1, b1);
  } else {
    mma(a0, b0);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int n = n0 + 16 * j + r;
      if (n >= N) continue;
      const float bn = EPI == EPI_MASK ? 0.f : bias[n];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int m = m0 + 16 * i + kq + e;
        if (m >= M) continue;
        const size_t o = (size_t)m * N + n;
        if (EPI == EPI_BIAS_RELU) Y[o] = fmaxf(acc[i][j][e] + bn, 0.f);
        if (EPI == EPI_BIAS) Y[o] = acc[i][j][e] + bn;
        if (EPI == EPI_MASK) Y[o] = A[o] > 0.f ? acc[i][j][e] : 0.f;
      }
    }
}

// Data gradient, for Cout % 16 == 0: dX[m][c] = sum over (tap, co) of bf16(dZ[(b, oy, ox)][co]) Wd[c][tap Cout + co]
// for the input pixel m = (b, iy, ix), over the taps whose output position reads it (oy stride =
// iy + 1 - ky, ox likewise, in range): dcol = dZ W followed by col2im, without the dcol matrix.
// At stride 2 the pixels run in four parity classes (blockIdx.z = 2 (iy % 2) + ix % 2): within a class
// the taps that reach a pixel are the same for every pixel (ky = 1 for even iy, ky = 0 and 2 for odd
// iy; kx likewise), so a wave runs 1, 2 or 4 of the 9 taps and no product of a tap that cannot reach
// it. Bp: Wd packed by k_pack_b (PACK_DGRAD, K = 9 Cout). Mc = the pixels of one class (all pixels at
// stride 1); Si even at stride 2. The layer's input is the previous layer's ReLU output A, so the
// epilogue applies that ReLU's backward as well: dX = 0 where A <= 0. A operand loads as k_conv_fwd.
template <int NT>
__global__ __launch_bounds__(256) void k_conv_dgrad(const float* __restrict__ dZ, int Si, int Cout, int So, int stride,
                                                    const bf16_t* __restrict__ Bp, const float* __restrict__ A,
                                                    float* __restrict__ dX, int Mc, int N) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m0 = xcd_block_x() * 256 + wave * 64, n0 = blockIdx.y * 16 * NT;
  const int r = lane & 15, kq = (lane >> 4) * 4;
  const int pl = lane >> 2, cq = (lane & 3) * 4, src = (4 * r + (lane >> 4)) * 4;
  // class geometry: pixel (b, q, p) of the class's Sq x Sq grid sits at (q step + py, p step + px)
  const int py = stride == 2 ? (int)(blockIdx.z >> 1) : 0, px = stride == 2 ? (int)(blockIdx.z & 1) : 0;
  const int Sq = Si / stride, SqSq = Sq * Sq;
  const int ky0 = stride == 2 ? 1 - py : 0, kyn = stride == 2 ? 2 + py : 3, kys = stride == 2 ? 2 : 1;
  const int kx0 = stride == 2 ? 1 - px : 0, kxn = stride == 2 ? 2 + px : 3, kxs = stride == 2 ? 2 : 1;
  int iy1[4], ix1[4], xo[4];
  const float* zb[4];
  bool mok[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + 16 * i + pl;
    mok[i] = m < Mc;
    const int mm = mok[i] ? m : 0;
    const int b = mm / SqSq, pix = mm - b * SqSq;
    const int q = pix / Sq, p = pix - q * Sq;
    const int iy = q * stride + py, ix = p * stride + px;
    iy1[i] = iy + 1;
    ix1[i] = ix + 1;
    zb[i] = dZ + (size_t)b * So * So * Cout + cq;
    xo[i] = mok[i] ? ((b * Si + iy) * Si + ix) * N : -1;
  }
  const int S = 9 * (Cout / 16);
  const bf16_t* bp = Bp + ((size_t)(n0 / 16) * S * 64 + lane) * 4;
  float4_t acc[4][NT];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
  const float4 zf4 = {0.f, 0.f, 0.f, 0.f};
  // the class's taps (nty x ntx of them) and their 16-channel k steps, flattened and pipelined as in
  // k_conv_fwd; ks is the step's index in the packed weights (tap (ky, kx) starts at (3 ky + kx) Cout / 16)
  const int ntx = (kxn - kx0 + kxs - 1) / kxs, ntaps = ((kyn - ky0 + kys - 1) / kys) * ntx;
  int tq = 0, c0 = 0, ks = 0;
  const float* ap[4];
  bool aok[4];
  auto setup = [&]() {
    const int ky = ky0 + kys * (tq / ntx), kx = kx0 + kxs * (tq % ntx);
    ks = (ky * 3 + kx) * (Cout / 16);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      // (iy + 1 - ky) is a multiple of the stride for every tap of the class
      const int ty = iy1[i] - ky, tx = ix1[i] - kx;
      const int oy = stride == 2 ? ty >> 1 : ty, ox = stride == 2 ? tx >> 1 : tx;
      aok[i] = mok[i] && ty >= 0 && tx >= 0 && oy < So && ox < So;
      // an output pixel clamped into the image: a valid address without a branch, its value then dropped
      ap[i] = zb[i] + (min(max(oy, 0), So - 1) * So + min(max(ox, 0), So - 1)) * Cout;
    }
  };
  auto load = [&](float4 (&ar)[4], short4_t (&br)[NT]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) ar[i] = load_or(reinterpret_cast<const float4*>(ap[i] + c0), aok[i], zf4);
#pragma unroll
    for (int j = 0; j < NT; ++j) br[j] = *reinterpret_cast<const short4_t*>(bp + ((size_t)j * S + ks) * 256);
    ++ks;
    c0 += 16;
    if (c0 == Cout) {  // next tap (past the last one the addresses are computed and never used)
      c0 = 0;
      ++tq;
      setup();
    }
  };
  auto mma = [&](const float4 (&ar)[4], const short4_t (&br)[NT]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const short4_t a = to_mfma_lanes(ar[i], src);
#pragma unroll
      for (int j = 0; j < NT; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, br[j], acc[i][j], 0, 0, 0);
    }
  };
  const int steps = ntaps * (Cout / 16);
  setup();
  float4 a0[4], a1[4];
  short4_t b0[NT], b1[NT];
  load(a0, b0);
  int s = 1;  // the next step to load; a0 holds step s - 1
  for (; s + 1 < steps; s += 2) {
    load(a1, b1);
    mma(a0, b0);
    load(a0, b0);
    mma(a1, b1);
  }
  if (s < steps) {
    load(a1, b1);
    mma(a0, b0);
    mma(a1, b1);
  } else {
    mma(a0, b0);
  }
  // D rows are the pixels m0 + 16 i + kq + e: their output offsets (-1 past Mc) come from lane
  // 4 (kq + e), which decoded that pixel
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int o = __shfl(xo[i], 4 * (kq + e));
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        const int n = n0 + 16 * j + r;
        if (o >= 0 && n < N) dX[(size_t)o + n] = A[(size_t)o + n] > 0.f ? acc[i][j][e] : 0.f;
      }
    }
}

// Weight gradient of a 3x3 convolution (TAPS = 9) or of the 1x1 head (TAPS = 1: the centre tap, stride
// 1) as an implicit product, for Cin % 16 == 0 and So % 16 == 0: parts[z][co][tap Cin + c] = sum over
// the output pixels p = (b, oy, ox) of split z of bf16(dZ[p][co]) bf16(X[b][oy stride + ky - 1]
// [ox stride + kx - 1][c]) (0 outside the image), with tap = 3 ky + kx (0 for the head): dW = dZ^T col
// without the im2col matrix or either transpose. dZ rows are ldz floats apart, ldz a multiple of
// 16 MT: Cout need not be a multiple of 16 when columns Cout .. ldz of dZ are zero (the head's padded
// dL/dlogits), the last tile then writing only its channels below Cout. Element offsets into X and
// dZ are 32-bit (the host checks both sizes). TAPS = 0 is the first layer (Cin = 1): the nine taps
// are the product's n (one 16-wide tile, lane r holding tap r < 9 and reading its own shifted pixels),
// parts[z][co][tap]. The bias gradient comes with it: the block of tap 0 and input-channel tile 0 of
// each (split, output-channel tile) also sums its FP32 dZ values per output channel (in segment,
// lane-group and wave order) into parts[z][Cout K + co], so a split's partials are the layer's
// [weights | biases], as in the flat parameter array. The product's k runs
// over pixels, in 16-pixel segments of one output row: a lane's four pixels are consecutive, so its
// four dZ values lie Cout floats apart and its four X values stride Cin floats apart, at constant
// offsets from one address (only the image's left and right edges mask a pixel). A block covers one
// tap, 16 MT output and 16 NT input channels of one split's rows; its four waves take every fourth
// segment and add their tiles through LDS in wave order (results do not depend on scheduling).
// Block b runs on XCD b % 8: the blocks of one split share an XCD, so the tiles that read the same
// pixels meet in one L2. nsplit % 8 == 0; ntiles = TAPS (Cin / 16 NT) ceil(Cout / 16 MT).
template <int MT, int NT, int TAPS>
__global__ __launch_bounds__(256) void k_conv_wgrad(const float* __restrict__ X, int Si, int Cin, int So, int stride,
                                                    const float* __restrict__ dZ, int ldz, int Cout, int rows,
                                                    int nsplit, int ntiles, float* __restrict__ parts) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int xcd = blockIdx.x & 7, jx = blockIdx.x >> 3;
  const int z = xcd * (nsplit >> 3) + jx / ntiles;
  int t = jx % ntiles;
  const int nc = TAPS == 0 ? 1 : Cin / (16 * NT);
  const int cm = TAPS == 0 ? t : t / (TAPS * nc);
  t -= TAPS == 0 ? t : cm * TAPS * nc;
  const int tapi = t / nc, cc = t - tapi * nc;
  const int tap = TAPS == 9 ? tapi : TAPS == 1 ? 4 : min(lane & 15, 8);
  const int ky = tap / 3, kx = tap - 3 * ky;
  const int co0 = cm * 16 * MT, ci0 = cc * 16 * NT;
  const int r = lane & 15, pq = (lane >> 4) * 4;
  const int spr = So >> 4;
  const bool bias_blk = tapi == 0 && cc == 0;  // uniform over the block
  float bsum[MT];
#pragma unroll
  for (int i = 0; i < MT; ++i) bsum[i] = 0.f;
  const int g_lo = (int)((long long)rows * z / nsplit) * spr, g_hi = (int)((long long)rows * (z + 1) / nsplit) * spr;
  float4_t acc[MT][NT];
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
  // loads of segment g into raw registers: dZ (MT x 4) and X (NT x 4), masked X values as 0
  auto load = [&](int g, float (&za)[MT][4], float (&xa)[NT][4]) {
    const int row = g / spr;
    const int ox = (g - row * spr) * 16 + pq;
    const int b = row / So, oy = row - b * So;
    const int zo = (row * So + ox) * ldz + co0 + r;
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e) za[i][e] = dZ[zo + e * ldz + 16 * i];
    const int iy = oy * stride + ky - 1;
    const int ix0 = ox * stride + kx - 1;
    if constexpr (TAPS == 0) {
      // lane r's tap: its row is its own (clamped into the image, the value then dropped)
      const bool rok = r < 9 && iy >= 0 && iy < Si;
      const int base = (b * Si + min(max(iy, 0), Si - 1)) * Si;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int ix = ix0 + e * stride;
        const float v = X[base + min(max(ix, 0), Si - 1)];
        xa[0][e] = rok && ix >= 0 && ix < Si ? v : 0.f;
      }
    } else if (iy >= 0 && iy < Si) {
      const int base = ((b * Si + iy) * Si + ix0) * Cin + ci0 + r;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int ix = ix0 + e * stride;
        const bool ok = ix >= 0 && ix < Si;
        const int xo = ok ? base + e * stride * Cin : 0;  // X itself when masked: loaded, then dropped
#pragma unroll
        for (int j = 0; j < NT; ++j) {
          const float v = X[xo + 16 * j];
          xa[j][e] = ok ? v : 0.f;
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < NT; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) xa[j][e] = 0.f;
    }
  };
  auto mfma = [&](const float (&za)[MT][4], const float (&xa)[NT][4]) {
    if (bias_blk) {
#pragma unroll
      for (int i = 0; i < MT; ++i) bsum[i] += ((za[i][0] + za[i][1]) + za[i][2]) + za[i][3];
    }
    short4_t a[MT], bq[NT];
#pragma unroll
    for (int i = 0; i < MT; ++i) a[i] = bf16x4_hw(za[i][0], za[i][1], za[i][2], za[i][3]);
#pragma unroll
    for (int j = 0; j < NT; ++j) bq[j] = bf16x4_hw(xa[j][0], xa[j][1], xa[j][2], xa[j][3]);
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
      for (int j = 0; j < NT; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], bq[j], acc[i][j], 0, 0, 0);
  };
  int g = g_lo + wave;
  for (; g + 4 < g_hi; g += 8) {
    float za0[MT][4], xa0[NT][4], za1[MT][4], xa1[NT][4];
    load(g, za0, xa0);
    load(g + 4, za1, xa1);
    mfma(za0, xa0);
    mfma(za1, xa1);
  }
  if (g < g_hi) {
    float za0[MT][4], xa0[NT][4];
    load(g, za0, xa0);
    mfma(za0, xa0);
  }
  // the bias sums of the four lane groups (the same value in every group: fixed pairing order)
  if (bias_blk) {
#pragma unroll
    for (int i = 0; i < MT; ++i) {
      bsum[i] += __shfl_xor(bsum[i], 16);
      bsum[i] += __shfl_xor(bsum[i], 32);
    }
  }
  // waves 1..3 hand their tiles to wave 0, which adds them in wave order and writes the partial
  __shared__ float4_t red[3][MT * NT][64];
  __shared__ float bred[3][MT][16];
  if (wave > 0) {
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
      for (int j = 0; j < NT; ++j) red[wave - 1][i * NT + j][lane] = acc[i][j];
    if (bias_blk && lane < 16) {
#pragma unroll
      for (int i = 0; i < MT; ++i) bred[wave - 1][i][lane] = bsum[i];
    }
  }
  __syncthreads();
  if (wave != 0) return;
  const int K = TAPS == 0 ? 9 : TAPS * Cin;
  float* const zpart = parts + (size_t)z * Cout * (K + 1);
  if (bias_blk && lane < 16) {
#pragma unroll
    for (int i = 0; i < MT; ++i) {
      const int co = co0 + 16 * i + r;
      const float s = ((bsum[i] + bred[0][i][r]) + bred[1][i][r]) + bred[2][i][r];
      if (co < Cout) zpart[(size_t)Cout * K + co] = s;
    }
  }
  float* out = zpart + (size_t)tapi * Cin;
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      float4_t s = acc[i][j];
#pragma unroll
      for (int w = 0; w < 3; ++w) s += red[w][i * NT + j][lane];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int co = co0 + 16 * i + pq + e;
        if (co < Cout && (TAPS != 0 || r < 9)) out[(size_t)co * K + ci0 + 16 * j + r] = s[e];
      }
    }
}

// out[m][n] = sum over z of parts[z][m][n] (+ bias, ReLU): a block covers 32 consecutive outputs
// with 8 interleaved z subsets each (z = q, q + 8, ...), added in subset order.
template <bool BIAS, bool RELU>
__global__ __launch_bounds__(256) void k_reduce_splits(const float* __restrict__ parts, int nsplit, long long M, int N,
                                                       long long split_stride, float* __restrict__ out,
                                                       long long ldo, const float* __restrict__ bias) {
  __shared__ float red[8][32];
  const long long total = M * N;
  const int o = threadIdx.x & 31, q = threadIdx.x >> 5;
  for (long long e0 = blockIdx.x * 32LL; e0 < total; e0 += (long long)gridDim.x * 32) {
    const long long e = e0 + o;
    float s0 = 0.f, s1 = 0.f;
    if (e < total) {
      int z = q;
      for (; z + 8 < nsplit; z += 16) {
        s0 += parts[z * split_stride + e];
        s1 += parts[(z + 8) * split_stride + e];
      }
      if (z < nsplit) s0 += parts[z * split_stride + e];
    }
    red[q][o] = s0 + s1;
    __syncthreads();
    if (q == 0 && e < total) {
      float s = 0.f;
#pragma unroll
      for (int k = 0; k < 8; ++k) s += red[k][o];
      const long long m = e / N;
      const int n = (int)(e - m * N);
      if (BIAS) s += bias[n];
      if (RELU) s = fmaxf(s, 0.f);
      out[m * ldo + n] = s;
    }
    __syncthreads();
  }
}

// col[row][k] (bf16, Kp per row): row = (b, oy, ox) of the So x So output, k = (ky * 3 + kx) * Cin + c
// reads X[b][oy * stride + ky - 1][ox * stride + kx - 1][c] (0 outside the image and for k >= 9 Cin).
// One thread per 4 consecutive k of a row (Kp % 4 == 0); when Cin % 4 == 0 the 4 share one tap and
// are one 16-byte load. rows * Kp / 4 < 2^31 (checked by the caller).
__global__ void k_im2col(const float* __restrict__ X, int Si, int Cin, int So, int stride, int rows, int Kp,
                         bf16_t* __restrict__ col) {
  const int Kq = Kp >> 2, K9 = 9 * Cin, SoSo = So * So;
  const unsigned total = (unsigned)rows * (unsigned)Kq;
  const bool vec = (Cin & 3) == 0;
  for (unsigned e = blockIdx.x * 256u + threadIdx.x; e < total; e += gridDim.x * 256u) {
    const int row = (int)(e / (unsigned)Kq);
    const int k0 = (int)(e - (unsigned)row * (unsigned)Kq) * 4;
    const int b = row / SoSo, pix = row - b * SoSo;
    const int oy = pix / So, ox = pix - oy * So;
    const float* xb = X + (size_t)b * Si * Si * Cin;
    float v[4];
    if (vec) {
      v[0] = v[1] = v[2] = v[3] = 0.f;
      if (k0 < K9) {
        const int tap = k0 / Cin, c = k0 - tap * Cin;
        const int ky = tap / 3, kx = tap - 3 * ky;
        const int iy = oy * stride + ky - 1, ix = ox * stride + kx - 1;
        if (iy >= 0 && iy < Si && ix >= 0 && ix < Si) {
          const float4 q = *reinterpret_cast<const float4*>(xb + ((size_t)iy * Si + ix) * Cin + c);
          v[0] = q.x;
          v[1] = q.y;
          v[2] = q.z;
          v[3] = q.w;
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int k = k0 + i;
        v[i] = 0.f;
        if (k < K9) {
          const int tap = k / Cin, c = k - tap * Cin;
          const int ky = tap / 3, kx = tap - 3 * ky;
          const int iy = oy * stride + ky - 1, ix = ox * stride + kx - 1;
          if (iy >= 0 && iy < Si && ix >= 0 && ix < Si) v[i] = xb[((size_t)iy * Si + ix) * Cin + c];
        }
      }
    }
    uint2 w;
    w.x = (unsigned)f2bf(v[0]) | ((unsigned)f2bf(v[1]) << 16);
    w.y = (unsigned)f2bf(v[2]) | ((unsigned)f2bf(v[3]) << 16);
    *reinterpret_cast<uint2*>(col + (size_t)row * Kp + k0) = w;
  }
}

// dX[b][iy][ix][c] = sum of dcol[(b, oy, ox)][(ky * 3 + kx) * Cin + c] over the taps that read it,
// in tap order. One thread per 4 consecutive channels of a pixel (Cin % 4 == 0, else one channel);
// pixels * Cin < 2^31 (checked by the caller).
__global__ void k_col2im(const float* __restrict__ dcol, int ldk, int Si, int Cin, int So, int stride, int total,
                         float* __restrict__ dX) {
  const int cw = (Cin & 3) == 0 ? 4 : 1;
  const int Cq = Cin / cw, SiSi = Si * Si;
  const unsigned n = (unsigned)(total / cw);
  for (unsigned e = blockIdx.x * 256u + threadIdx.x; e < n; e += gridDim.x * 256u) {
    const int p = (int)(e / (unsigned)Cq);
    const int c = (int)(e - (unsigned)p * (unsigned)Cq) * cw;
    const int b = p / SiSi, pix = p - b * SiSi;
    const int iy = pix / Si, ix = pix - iy * Si;
    float4 s = {0.f, 0.f, 0.f, 0.f};
    for (int ky = 0; ky < 3; ++ky) {
      const int ty = iy + 1 - ky;
      if (ty < 0 || ty % stride) continue;
      const int oy = ty / stride;
      if (oy >= So) continue;
      for (int kx = 0; kx < 3; ++kx) {
        const int tx = ix + 1 - kx;
        if (tx < 0 || tx % stride) continue;
        const int ox = tx / stride;
        if (ox >= So) continue;
        const float* src = dcol + ((size_t)(b * So + oy) * So + ox) * ldk + (ky * 3 + kx) * Cin + c;
        if (cw == 4) {
          const float4 q = *reinterpret_cast<const float4*>(src);
          s.x += q.x;
          s.y += q.y;
          s.z += q.z;
          s.w += q.w;
        } else {
          s.x += *src;
        }
      }
    }
    float* dst = dX + (size_t)p * Cin + c;
    if (cw == 4)
      *reinterpret_cast<float4*>(dst) = s;
    else
      *dst = s.x;
  }
}

// T[c][r] = bf16(X[r][c]) for c < Cp, r < Rp (0 outside R x C); 32 x 32 tiles through LDS
template <class TIn>
__global__ __launch_bounds__(256) void k_transpose_bf16(const TIn* __restrict__ X, long long R, int C, long long ldx,
                                                        bf16_t* __restrict__ T, long long ldt, int Cp, long long Rp) {
  __shared__ float tile[32][33];
  const long long r0 = (long long)blockIdx.x * 32;
  const int c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int j = ty; j < 32; j += 8) {
    const long long r = r0 + j;
    const int c = c0 + tx;
    tile[j][tx] = (r < R && c < C) ? load_f(X + r * ldx + c) : 0.f;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int c = c0 + j;
    const long long r = r0 + tx;
    if (c < Cp && r < Rp) T[(long long)c * ldt + r] = f2bf(tile[tx][j]);
  }
}

// Y[r][k] = bf16(X[r][k]) for k < C, 0 for C <= k < Cp
__global__ void k_to_bf16_pad(const float* __restrict__ X, long long R, int C, long long ldx, bf16_t* __restrict__ Y,
                              int Cp) {
  const long long total = R * Cp;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
    const long long r = e / Cp;
    const int k = (int)(e % Cp);
    Y[e] = f2bf(k < C ? X[r * ldx + k] : 0.f);
  }
}

// dA *= (A > 0): the ReLU's backward from its output
__global__ void k_relu_bwd(float* __restrict__ dA, const float* __restrict__ A, long long n) {
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < n; e += (long long)gridDim.x * 256)
    if (!(A[e] > 0.f)) dA[e] = 0.f;
}

// Column sums of X[R][C] over row chunks of `chunk` rows: part[chunk index][c]. Threads split the
// chunk's rows into 256 / Cg interleaved subsets per column group of Cg = min(C, 256) columns, each
// thread with four accumulators over every fourth row of its subset, added in a fixed order.
__global__ __launch_bounds__(256) void k_colsum_part(const float* __restrict__ X, long long R, int C, long long ldx,
                                                     long long chunk, float* __restrict__ part) {
  __shared__ float red[256];
  const int Cg = min(C, 256), per = 256 / Cg;
  const int t = threadIdx.x, c_off = t % Cg, sub = t / Cg;
  const long long r_lo = blockIdx.x * chunk, r_hi = min(R, r_lo + chunk);
  for (int cb = 0; cb < C; cb += Cg) {
    const int c = cb + c_off;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    if (sub < per && c < C) {
      long long r = r_lo + sub;
      for (; r + 3 * per < r_hi; r += 4 * per) {
        s0 += X[r * ldx + c];
        s1 += X[(r + per) * ldx + c];
        s2 += X[(r + 2 * per) * ldx + c];
        s3 += X[(r + 3 * per) * ldx + c];
      }
      for (; r < r_hi; r += per) s0 += X[r * ldx + c];
    }
    red[t] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (t < Cg && c < C) {
      float a = 0.f;
      for (int q = 0; q < per; ++q) a += red[q * Cg + t];
      part[(long long)blockIdx.x * C + c] = a;
    }
    __syncthreads();
  }
}
// out[c] = sum over q of part[q][c]: one block per column, 256 interleaved subsets then a tree
__global__ __launch_bounds__(256) void k_colsum_final(const float* __restrict__ part, int nparts, int C,
                                                      float* __restrict__ out) {
  __shared__ float red[256];
  const int c = blockIdx.x, t = threadIdx.x;
  float s = 0.f;
  for (int q = t; q < nparts; q += 256) s += part[(long long)q * C + c];
  red[t] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (t < k) red[t] += red[t + k];
    __syncthreads();
  }
  if (t == 0) out[c] = red[0];
}

// Per-image standardisation, (x - mean) / std over the S x S pixels with FP64 sums: part[b][p] holds
// the sums of the p-th of STD_PARTS contiguous segments of image b (blockIdx (p, b)), then
// k_std_apply adds them in order and normalises.
constexpr int STD_PARTS = 64;
__global__ __launch_bounds__(256) void k_std_part(const float* __restrict__ X, long long npx,
                                                  double* __restrict__ part) {
  __shared__ double s1[256], s2[256];
  const int p = blockIdx.x, b = blockIdx.y;
  const long long seg = (npx + STD_PARTS - 1) / STD_PARTS;
  const long long lo = p * seg, hi = min(npx, lo + seg);
  const float* x = X + b * npx;
  double a = 0.0, q = 0.0;
  for (long long e = lo + threadIdx.x; e < hi; e += 256) {
    const double v = x[e];
    a += v;
    q += v * v;
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = q;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      s1[threadIdx.x] += s1[threadIdx.x + k];
      s2[threadIdx.x] += s2[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[((long long)b * STD_PARTS + p) * 2] = s1[0];
    part[((long long)b * STD_PARTS + p) * 2 + 1] = s2[0];
  }
}
// Y = (X - mean) / std per image; X and Y may be the same buffer.
__global__ __launch_bounds__(256) void k_std_apply(const float* X, float* Y, long long npx,
                                                   const double* __restrict__ part) {
  __shared__ float mi[2];
  const int b = blockIdx.y;
  if (threadIdx.x == 0) {
    double a = 0.0, q = 0.0;
    for (int p = 0; p < STD_PARTS; ++p) {
      a += part[((long long)b * STD_PARTS + p) * 2];
      q += part[((long long)b * STD_PARTS + p) * 2 + 1];
    }
    const double mean = a / (double)npx;
    const double var = fmax(q / (double)npx - mean * mean, 1e-30);
    mi[0] = (float)mean;
    mi[1] = (float)(1.0 / sqrt(var));
  }
  __syncthreads();
  const float m = mi[0], inv = mi[1];
  const float* x = X + b * npx;
  float* y = Y + b * npx;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < npx; e += (long long)gridDim.x * 256)
    y[e] = (x[e] - m) * inv;
}

__device__ inline float softplus(float x) { return fmaxf(x, 0.f) + log1pf(expf(-fabsf(x))); }
__device__ inline float sigmoid(float x) { return 1.f / (1.f + expf(-x)); }

// One wave per cell: binary cross-entropy of the location logit (positives weighted by pos_w) and
// the softmax cross-entropy of the orientation logits when the cell has a bin. Gradients are
// scaled by heat_scale and bin_scale (the means' divisors); cell_loss receives the unscaled terms.
__global__ __launch_bounds__(64) void k_loss(const float* __restrict__ logits, int Nh, const float* __restrict__ heat_t,
                                             const int* __restrict__ bin_t, float pos_w, float heat_scale,
                                             float bin_scale, float* __restrict__ dlog, int ldd,
                                             float* __restrict__ cell_loss) {
  const long long cell = blockIdx.x;
  const float* o = logits + cell * Nh;
  float* d = dlog + cell * ldd;
  const int lane = threadIdx.x;
  for (int k = Nh + lane; k < ldd; k += 64) d[k] = 0.f;  // the row's padding
  const int nb = Nh - 1;
  const int tb = bin_t[cell] < nb ? bin_t[cell] : -1;  // a bin past the last counts as none
  if (lane == 0) {
    const float z = o[0], y = heat_t[cell];
    cell_loss[2 * cell] = pos_w * y * softplus(-z) + (1.f - y) * softplus(z);
    d[0] = heat_scale * (pos_w * y * (sigmoid(z) - 1.f) + (1.f - y) * sigmoid(z));
  }
  if (tb < 0) {
    for (int k = lane; k < nb; k += 64) d[1 + k] = 0.f;
    if (lane == 0) cell_loss[2 * cell + 1] = 0.f;
    return;
  }
  float mx = -INFINITY;
  for (int k = lane; k < nb; k += 64) mx = fmaxf(mx, o[1 + k]);
  for (int s = 32; s > 0; s >>= 1) mx = fmaxf(mx, __shfl_xor(mx, s));
  float se = 0.f;
  for (int k = lane; k < nb; k += 64) se += expf(o[1 + k] - mx);
  for (int s = 32; s > 0; s >>= 1) se += __shfl_xor(se, s);
  const float lse = mx + logf(se);
  for (int k = lane; k < nb; k += 64) d[1 + k] = bin_scale * (expf(o[1 + k] - lse) - (k == tb ? 1.f : 0.f));
  if (lane == 0) cell_loss[2 * cell + 1] = lse - o[1 + tb];
}

// out[c] = logits[c][0] (the location logit of every cell)
__global__ void k_column0(const float* __restrict__ logits, int Nh, long long cells, float* __restrict__ out) {
  for (long long c = blockIdx.x * 256LL + threadIdx.x; c < cells; c += (long long)gridDim.x * 256)
    out[c] = logits[c * Nh];
}

// out[q][*] = logits[cells[q]][*] (the logits of chosen cells)
__global__ void k_gather_cells(const float* __restrict__ logits, int Nh, const int* __restrict__ cells, int n,
                               float* __restrict__ out) {
  const int q = blockIdx.x;
  if (q >= n) return;
  const float* src = logits + (long long)cells[q] * Nh;
  for (int k = threadIdx.x; k < Nh; k += blockDim.x) out[(long long)q * Nh + k] = src[k];
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       long long n, float lr_t, float b1, float b2, float eps) {
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < n; e += (long long)gridDim.x * 256) {
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.f - b1) * gi;
    const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    w[e] -= lr_t * mi / (sqrtf(vi) + eps);
  }
}

// ------------------------------------------------------------------ host
template <class T>
struct Buf {
  T* p = nullptr;
  size_t n = 0;
  void ensure(size_t count) {
    if (count <= n) return;
    release();
    CNN_HIP(hipMalloc((void**)&p, count * sizeof(T)));
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~Buf() { release(); }
};

struct Layer {
  int cin, cout, stride, si, so;  // channels, stride, input and output edge
  int K, Kp;                      // 9 cin (1 cin for the head) and its multiple of 16
  long long w_off, b_off;         // offsets in the flat parameter array
};

// C[M][N] (ldc) = A[M][lda] . B[N][ldb]^T (+ bias, ReLU), split over K when the output tile grid is
// too small to fill the GPU.
template <int WN>
void gemm_wn(hipStream_t st, Buf<float>& parts, const bf16_t* A, int lda, const bf16_t* B, int ldb, float* C,
             long long ldc, long long M, int N, int K, const float* bias, bool relu) {
  constexpr int BM = 128 / WN, BN = 32 * WN;
  const long long gm = (M + BM - 1) / BM;
  const int gn = (N + BN - 1) / BN;
  CNN_REQUIRE(gm < (1LL << 31) && gn <= 65535, TM_ERR_INVALID, "gemm: output too large");
  const long long tiles = gm * gn;
  int nsplit = 1;
  if (tiles < 1024 && K >= 2048) {
    // about 2048 blocks, at least 2048 of K per block, partials <= 256 MB
    nsplit = (int)std::min<long long>(K / 2048, std::max<long long>(1, 2048 / tiles));
    const long long cap = (64LL << 20) / std::max<long long>(1, M * N);
    nsplit = (int)std::max<long long>(1, std::min<long long>(nsplit, cap));
  }
  if (nsplit <= 1) {
    dim3 g((unsigned)gm, gn, 1);
    if (bias && relu)
      hipLaunchKernelGGL((k_gemm_nt<true, true, WN>), g, dim3(256), 0, st, A, lda, B, ldb, C, ldc, M, N, K, K, 0LL,
                         bias);
    else if (bias)
      hipLaunchKernelGGL((k_gemm_nt<true, false, WN>), g, dim3(256), 0, st, A, lda, B, ldb, C, ldc, M, N, K, K, 0LL,
                         bias);
    else
      hipLaunchKernelGGL((k_gemm_nt<false, false, WN>), g, dim3(256), 0, st, A, lda, B, ldb, C, ldc, M, N, K, K, 0LL,
                         nullptr);
    check_launch("k_gemm_nt");
    return;
  }
  const int kchunk = (K / 16 + nsplit - 1) / nsplit * 16;
  nsplit = (K + kchunk - 1) / kchunk;
  const long long ss = M * N;
  parts.ensure((size_t)ss * nsplit);
  hipLaunchKernelGGL((k_gemm_nt<false, false, WN>), dim3((unsigned)gm, gn, nsplit), dim3(256), 0, st, A, lda, B, ldb,
                     parts.p, (long long)N, M, N, K, kchunk, ss, nullptr);
  check_launch("k_gemm_nt split");
  const unsigned g = grid1((M * N + 31) / 32, 1, 1 << 16);
  if (bias && relu)
    hipLaunchKernelGGL((k_reduce_splits<true, true>), dim3(g), dim3(256), 0, st, parts.p, nsplit, M, N, ss, C, ldc, bias);
  else if (bias)
    hipLaunchKernelGGL((k_reduce_splits<true, false>), dim3(g), dim3(256), 0, st, parts.p, nsplit, M, N, ss, C, ldc, bias);
  else
    hipLaunchKernelGGL((k_reduce_splits<false, false>), dim3(g), dim3(256), 0, st, parts.p, nsplit, M, N, ss, C, ldc,
                       nullptr);
  check_launch("k_reduce_splits");
}

// C = A . B^T on k_gemm_nt4's 128 x 128 blocks, split over K into fixed-order partials when the tile
// grid is too small to fill the GPU (about 2048 blocks, at least 512 of K per block).
void gemm4(hipStream_t st, Buf<float>& parts, const bf16_t* A, int lda, const bf16_t* B, int ldb, float* C,
           long long ldc, long long M, int N, int K, const float* bias, bool relu) {
  const long long gm = (M + 127) / 128;
  const int gn = (N + 127) / 128;
  CNN_REQUIRE(gm < (1LL << 31) && gn <= 65535, TM_ERR_INVALID, "gemm: output too large");
  const long long tiles = gm * gn;
  int nsplit = 1;
  if (tiles < 1024 && K >= 1024) {
    nsplit = (int)std::min<long long>(K / 512, std::max<long long>(1, 2048 / tiles));
    const long long cap = (64LL << 20) / std::max<long long>(1, M * N);
    nsplit = (int)std::max<long long>(1, std::min<long long>(nsplit, cap));
  }
  if (nsplit <= 1) {
    const dim3 g((unsigned)gm, gn, 1);
    if (bias && relu)
      hipLaunchKernelGGL((k_gemm_nt4<true, true>), g, dim3(256), 0, st, A, lda, B, ldb, C, ldc, M, N, K, K, 0LL, bias);
    else if (bias)
      hipLaunchKernelGGL((k_gemm_nt4<true, false>), g, dim3(256), 0, st, A, lda, B, ldb, C, ldc, M, N, K, K, 0LL, bias);
    else
      hipLaunchKernelGGL((k_gemm_nt4<false, false>), g, dim3(256), 0, st, A, lda, B, ldb, C, ldc, M, N, K, K, 0LL,
                         nullptr);
    check_launch("k_gemm_nt4");
    return;
  }
  const int kchunk = (K / 16 + nsplit - 1) / nsplit * 16;
  nsplit = (K + kchunk - 1) / kchunk;
  const long long ss = M * N;
  parts.ensure((size_t)ss * nsplit);
  hipLaunchKernelGGL((k_gemm_nt4<false, false>), dim3((unsigned)gm, gn, nsplit), dim3(256), 0, st, A, lda, B, ldb,
                     parts.p, (long long)N, M, N, K, kchunk, ss, nullptr);
  check_launch("k_gemm_nt4 split");
  const unsigned g = grid1((M * N + 31) / 32, 1, 1 << 16);
  if (bias && relu)
    hipLaunchKernelGGL((k_reduce_splits<true, true>), dim3(g), dim3(256), 0, st, parts.p, nsplit, M, N, ss, C, ldc, bias);
  else if (bias)
    hipLaunchKernelGGL((k_reduce_splits<true, false>), dim3(g), dim3(256), 0, st, parts.p, nsplit, M, N, ss, C, ldc, bias);
  else
    hipLaunchKernelGGL((k_reduce_splits<false, false>), dim3(g), dim3(256), 0, st, parts.p, nsplit, M, N, ss, C, ldc,
                       nullptr);
  check_launch("k_reduce_splits");
}

// Block shape by operand size: 128 x 128 (k_gemm_nt4) when M and N are both 128 or more and the rows
// allow 16-byte loads; else 128 x 32 when N <= 32, 32 x 128 when M <= 32, 64 x 64 otherwise.
void gemm(hipStream_t st, Buf<float>& parts, const bf16_t* A, int lda, const bf16_t* B, int ldb, float* C,
          long long ldc, long long M, int N, int K, const float* bias, bool relu) {
  CNN_REQUIRE(K % 16 == 0 && lda % 4 == 0 && ldb % 4 == 0, TM_ERR_INVALID, "gemm: K, lda, ldb not padded");
  if (M >= 128 && N >= 128 && lda % 8 == 0 && ldb % 8 == 0)
    gemm4(st, parts, A, lda, B, ldb, C, ldc, M, N, K, bias, relu);
  else if (N <= 32)
    gemm_wn<1>(st, parts, A, lda, B, ldb, C, ldc, M, N, K, bias, relu);
  else if (M <= 32)
    gemm_wn<4>(st, parts, A, lda, B, ldb, C, ldc, M, N, K, bias, relu);
  else
    gemm_wn<2>(st, parts, A, lda, B, ldb, C, ldc, M, N, K, bias, relu);
}

// Output channels per wave tile: 16 NT, NT = 4 from 64 channels up.
inline int conv_nt(int n) { return n >= 64 ? 4 : (n + 15) / 16; }
// 16-row blocks of a packed operand with N rows: a multiple of 4, so a wave's tiles stay inside.
inline int packed_blocks(int N) { return ((N + 15) / 16 + 3) / 4 * 4; }
inline size_t packed_size(int N, int K) { return (size_t)packed_blocks(N) * 16 * K; }
// B packed in MFMA operand order (k_pack_b) into Bp, packed_size(N, K) elements, K % 16 == 0.
template <int MODE>
void pack_b(hipStream_t st, const float* W, int N, int K, int ldw, int Kw, int Cin, int Cout, bf16_t* Bp) {
  hipLaunchKernelGGL((k_pack_b<MODE>), dim3(grid1((long long)packed_size(N, K))), dim3(256), 0, st, W, N, K, ldw, Kw,
                     Cin, Cout, packed_blocks(N), Bp);
  check_launch("k_pack_b");
}
template <int NT, int TAPS, int EPI>
void conv_fwd_nt(hipStream_t st, const float* X, int Si, int Cin, int So, int stride, const bf16_t* Bp, float* Y,
                 int M, int N, const float* bias, const float* A) {
  const dim3 g((unsigned)((M + 255) / 256), (unsigned)((N + 16 * NT - 1) / (16 * NT)));
  hipLaunchKernelGGL((k_conv_fwd<NT, TAPS, EPI>), g, dim3(256), 0, st, X, Si, Cin, So, stride, Bp, Y, M, N, bias, A);
  check_launch("k_conv_fwd");
}
// A 3x3 (TAPS 9) or 1x1 (TAPS 1) layer as an implicit product without an im2col matrix (Cin % 16 == 0),
// B packed by pack_b<PACK_ROWS or PACK_COLS>(N rows, K = TAPS Cin), with the epilogue EPI (bias and
// ReLU, bias, or the mask of A's ReLU).
template <int TAPS, int EPI>
void conv_fwd(hipStream_t st, const float* X, int Si, int Cin, int So, int stride, const bf16_t* Bp, float* Y,
              long long M, int N, const float* bias, const float* A = nullptr) {
  CNN_REQUIRE(Cin % 16 == 0, TM_ERR_INVALID, "conv: Cin must be a multiple of 16");
  CNN_REQUIRE((long long)Si * Si * Cin < (1LL << 31) && M * N < (1LL << 31), TM_ERR_INVALID, "conv: image too large");
  switch (conv_nt(N)) {
    case 1: conv_fwd_nt<1, TAPS, EPI>(st, X, Si, Cin, So, stride, Bp, Y, (int)M, N, bias, A); break;
    case 2: conv_fwd_nt<2, TAPS, EPI>(st, X, Si, Cin, So, stride, Bp, Y, (int)M, N, bias, A); break;
    case 3: conv_fwd_nt<3, TAPS, EPI>(st, X, Si, Cin, So, stride, Bp, Y, (int)M, N, bias, A); break;
    default: conv_fwd_nt<4, TAPS, EPI>(st, X, Si, Cin, So, stride, Bp, Y, (int)M, N, bias, A); break;
  }
}

// Data gradient of a 3x3 layer into dX[B][Si][Si][Cin] (k_conv_dgrad; Cout % 16 == 0), its weights
// packed by pack_b<PACK_DGRAD>, masked by the ReLU that produced the layer's input A. At stride 2 one
// grid layer per pixel-parity class.
template <int NT>
void conv_dgrad_nt(hipStream_t st, const float* dZ, const Layer& L, int batch, const bf16_t* Bp, const float* A,
                   float* dX) {
  const int Sq = L.si / L.stride;
  const int Mc = batch * Sq * Sq;
  const dim3 g((unsigned)((Mc + 255) / 256), (unsigned)((L.cin + 16 * NT - 1) / (16 * NT)), L.stride == 2 ? 4 : 1);
  hipLaunchKernelGGL((k_conv_dgrad<NT>), g, dim3(256), 0, st, dZ, L.si, L.cout, L.so, L.stride, Bp, A, dX, Mc, L.cin);
  check_launch("k_conv_dgrad");
}
void conv_dgrad(hipStream_t st, const float* dZ, const Layer& L, int batch, const bf16_t* Bp, const float* A,
                float* dX) {
  CNN_REQUIRE(L.cout % 16 == 0 && L.si % L.stride == 0, TM_ERR_INVALID, "dgrad: Cout must be a multiple of 16");
  CNN_REQUIRE((long long)L.so * L.so * L.cout < (1LL << 31) && (long long)batch * L.si * L.si * L.cin < (1LL << 31),
              TM_ERR_INVALID, "dgrad: layer too large");
  switch (conv_nt(L.cin)) {
    case 1: conv_dgrad_nt<1>(st, dZ, L, batch, Bp, A, dX); break;
    case 2: conv_dgrad_nt<2>(st, dZ, L, batch, Bp, A, dX); break;
    case 3: conv_dgrad_nt<3>(st, dZ, L, batch, Bp, A, dX); break;
    default: conv_dgrad_nt<4>(st, dZ, L, batch, Bp, A, dX); break;
  }
}

void transpose_bf16(hipStream_t st, const float* X, long long R, int C, long long ldx, bf16_t* T, long long ldt, int Cp,
                    long long Rp) {
  hipLaunchKernelGGL(k_transpose_bf16<float>, dim3((unsigned)((Rp + 31) / 32), (Cp + 31) / 32), dim3(256), 0, st, X, R,
                     C, ldx, T, ldt, Cp, Rp);
  check_launch("k_transpose_bf16");
}
void transpose_bf16(hipStream_t st, const bf16_t* X, long long R, int C, long long ldx, bf16_t* T, long long ldt, int Cp,
                    long long Rp) {
  hipLaunchKernelGGL(k_transpose_bf16<bf16_t>, dim3((unsigned)((Rp + 31) / 32), (Cp + 31) / 32), dim3(256), 0, st, X, R,
                     C, ldx, T, ldt, Cp, Rp);
  check_launch("k_transpose_bf16");
}
void to_bf16_pad(hipStream_t st, const float* X, long long R, int C, long long ldx, bf16_t* Y, int Cp) {
  hipLaunchKernelGGL(k_to_bf16_pad, dim3(grid1(R * Cp)), dim3(256), 0, st, X, R, C, ldx, Y, Cp);
  check_launch("k_to_bf16_pad");
}
// Column sums of X[R][C] (rows ldx apart) in two fixed-order passes: about 2048 row chunks, then one
// block per column.
void colsum(hipStream_t st, Buf<float>& part, const float* X, long long R, int C, long long ldx, float* out) {
  const long long per = 256 / std::min(C, 256);
  const long long chunk = std::max<long long>(4 * per, (R + 2047) / 2048);
  const int np = (int)((R + chunk - 1) / chunk);
  part.ensure((size_t)np * C);
  hipLaunchKernelGGL(k_colsum_part, dim3(np), dim3(256), 0, st, X, R, C, ldx, chunk, part.p);
  check_launch("k_colsum_part");
  hipLaunchKernelGGL(k_colsum_final, dim3(C), dim3(256), 0, st, part.p, np, C, out);
  check_launch("k_colsum_final");
}

// dW[cout][9 cin] of a 3x3 layer and its bias gradient (written after it: dW points at the layer's
// weights in the flat gradient, the biases follow them) straight from its input X and output gradient
// dZ (k_conv_wgrad), the pixel splits' partials added in split order: about 2048 blocks, a split holding at least one
// output row, the partials at most 64 M floats.
template <int MT, int NT, int TAPS>
void conv_wgrad_t(hipStream_t st, Buf<float>& parts, const float* X, const Layer& L, int rows, const float* dZ,
                  int ldz, float* dW) {
  const int ntiles = (TAPS == 0 ? 1 : TAPS * (L.cin / (16 * NT))) * ((L.cout + 16 * MT - 1) / (16 * MT));
  const long long per = (long long)L.cout * (L.K + 1);  // weights, then biases
  long long ns = std::max<long long>(1, 2048 / ntiles);
  ns = std::min<long long>(ns, rows);
  ns = std::min<long long>(ns, (64LL << 20) / per);
  const int nsplit = (int)std::max<long long>(8, ns / 8 * 8);
  CNN_REQUIRE((long long)nsplit * ntiles < (1LL << 31), TM_ERR_INVALID, "wgrad: grid too large");
  parts.ensure((size_t)nsplit * per);
  hipLaunchKernelGGL((k_conv_wgrad<MT, NT, TAPS>), dim3((unsigned)(nsplit * ntiles)), dim3(256), 0, st, X, L.si,
                     L.cin, L.so, L.stride, dZ, ldz, L.cout, rows, nsplit, ntiles, parts.p);
  check_launch("k_conv_wgrad");
  const unsigned g = grid1((per + 31) / 32, 1, 1 << 16);
  hipLaunchKernelGGL((k_reduce_splits<false, false>), dim3(g), dim3(256), 0, st, parts.p, nsplit, 1LL, (int)per, per,
                     dW, per, nullptr);
  check_launch("k_reduce_splits");
}
template <int MT, int TAPS>
void conv_wgrad_m(hipStream_t st, Buf<float>& parts, const float* X, const Layer& L, int rows, const float* dZ,
                  int ldz, float* dW) {
  const int q = L.cin / 16;
  if constexpr (TAPS == 0)
    conv_wgrad_t<MT, 1, 0>(st, parts, X, L, rows, dZ, ldz, dW);
  else if (q % 4 == 0)
    conv_wgrad_t<MT, 4, TAPS>(st, parts, X, L, rows, dZ, ldz, dW);
  else if (q % 3 == 0)
    conv_wgrad_t<MT, 3, TAPS>(st, parts, X, L, rows, dZ, ldz, dW);
  else if (q % 2 == 0)
    conv_wgrad_t<MT, 2, TAPS>(st, parts, X, L, rows, dZ, ldz, dW);
  else
    conv_wgrad_t<MT, 1, TAPS>(st, parts, X, L, rows, dZ, ldz, dW);
}
// Whether k_conv_wgrad applies: input channels in multiples of 16 (and output channels for a 3x3
// layer), output rows in 16-pixel segments, X and dZ (ldz floats per row) under 2^31 elements.
bool wgrad_implicit(const Layer& L, int batch, int ldz, bool head) {
  const long long M = (long long)batch * L.so * L.so;
  return (L.cin % 16 == 0 || (L.cin == 1 && !head)) && (head || L.cout % 16 == 0) && ldz % 16 == 0 &&
         L.so % 16 == 0 && M * ldz < (1LL << 31) && (long long)batch * L.si * L.si * L.cin < (1LL << 31);
}
// dZ rows ldz floats apart (ldz a multiple of 16, zeros past cout); tiles of 16 MT output channels
// with MT dividing ldz / 16, so that no tile reads past a row.
template <int TAPS>
void conv_wgrad(hipStream_t st, Buf<float>& parts, const float* X, const Layer& L, int batch, const float* dZ, int ldz,
                float* dW) {
  const int rows = batch * L.so, q = ldz / 16;
  if (q % 4 == 0)
    conv_wgrad_m<4, TAPS>(st, parts, X, L, rows, dZ, ldz, dW);
  else if (q % 3 == 0)
    conv_wgrad_m<3, TAPS>(st, parts, X, L, rows, dZ, ldz, dW);
  else if (q % 2 == 0)
    conv_wgrad_m<2, TAPS>(st, parts, X, L, rows, dZ, ldz, dW);
  else
    conv_wgrad_m<1, TAPS>(st, parts, X, L, rows, dZ, ldz, dW);
}

void im2col(hipStream_t st, const float* X, const Layer& L, long long M, bf16_t* col) {
  CNN_REQUIRE(M * (L.Kp / 4) < (1LL << 31), TM_ERR_INVALID, "im2col: layer too large for one launch");
  hipLaunchKernelGGL(k_im2col, dim3(grid1(M * (L.Kp / 4))), dim3(256), 0, st, X, L.si, L.cin, L.so, L.stride, (int)M,
                     L.Kp, col);
  check_launch("k_im2col");
}

}  // namespace

struct tm_cnn {
  int device = 0;
  hipStream_t stream = nullptr;
  tm_cnn_config cfg{};
  std::vector<Layer> layers;  // conv layers, then the head (a 1x1 "layer" with K = C_L)
  int G = 0, Nh = 0;
  long long n_params = 0;
  int batch = 0;  // images of the last forward pass
  long long adam_t = 0;
  Buf<float> w, g, m, v;          // parameters, gradients, Adam moments (flat)
  std::vector<Buf<float>> act;    // act[0]: standardised input; act[l + 1]: layer l output; last: logits
  Buf<bf16_t> col, colT, wb, wbT, dzb, dzT;
  Buf<float> dA0, dA1, dcol, parts, csum, dlog, cell_loss, heat_t;
  Buf<int> bin_t, cells;
  Buf<float> gathered;
  Buf<double> std_part;
  std::vector<float> h_loss;
  long long loss_cells = 0, loss_bins = 0;  // cells and labelled cells of the last backward pass

  ~tm_cnn() {
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
  }
};

namespace {

void require(tm_cnn* h) {
  CNN_REQUIRE(h, TM_ERR_INVALID, "handle is NULL");
  CNN_HIP(hipSetDevice(h->device));
}

long long rows_of(const tm_cnn* h, const Layer& L) { return (long long)h->batch * L.so * L.so; }

// forward of layer l from act[l] into act[l + 1]
void forward_layer(tm_cnn* h, int l) {
  const Layer& L = h->layers[l];
  hipStream_t st = h->stream;
  const bool head = l == (int)h->layers.size() - 1;
  const long long M = rows_of(h, L);
  if (L.cin % 16 == 0) {
    pack_b<PACK_ROWS>(st, h->w.p + L.w_off, L.cout, L.K, L.K, L.K, 0, 0, h->wb.p);
    if (head)
      conv_fwd<1, EPI_BIAS>(st, h->act[l].p, L.si, L.cin, L.so, 1, h->wb.p, h->act[l + 1].p, M, L.cout,
                            h->w.p + L.b_off);
    else
      conv_fwd<9, EPI_BIAS_RELU>(st, h->act[l].p, L.si, L.cin, L.so, L.stride, h->wb.p, h->act[l + 1].p, M, L.cout,
                                 h->w.p + L.b_off);
    return;
  }
  to_bf16_pad(st, h->w.p + L.w_off, L.cout, L.K, L.K, h->wb.p, L.Kp);
  if (head)
    to_bf16_pad(st, h->act[l].p, M, L.cin, L.cin, h->col.p, L.Kp);
  else
    im2col(st, h->act[l].p, L, M, h->col.p);
  gemm(st, h->parts, h->col.p, L.Kp, h->wb.p, L.Kp, h->act[l + 1].p, L.cout, M, L.cout, L.Kp, h->w.p + L.b_off, !head);
}

void alloc_for_batch(tm_cnn* h, int batch) {
  h->batch = batch;
  size_t col = 0, colT = 0, dzT = 0, dzb = 0, grad = 0, dcol = 0, wmax = 0;
  for (size_t l = 0; l < h->layers.size(); ++l) {
    const Layer& L = h->layers[l];
    const long long M = (long long)batch * L.so * L.so, Mp = round16ll(M);
    col = std::max(col, (size_t)(M * L.Kp));
    colT = std::max(colT, (size_t)(Mp * L.Kp));
    dzT = std::max(dzT, (size_t)(Mp * L.cout));
    dzb = std::max(dzb, (size_t)(M * round16(L.cout)));
    grad = std::max(grad, (size_t)(M * L.cout));
    grad = std::max(grad, (size_t)((long long)batch * L.si * L.si * L.cin));
    dcol = std::max(dcol, (size_t)(M * L.K));
    wmax = std::max(wmax, (size_t)round16(L.cout) * L.Kp);
    // packed operands: the forward's weights, the data gradient's (3x3: 9 Cout per input channel;
    // head: its padded outputs per input channel)
    wmax = std::max(wmax, packed_size(L.cout, L.Kp));
    wmax = std::max(wmax, packed_size(L.cin, 9 * round16(L.cout)));
    h->act[l + 1].ensure((size_t)(M * L.cout));
  }
  const Layer& L0 = h->layers[0];
  h->act[0].ensure((size_t)batch * L0.si * L0.si);
  h->col.ensure(col);
  h->colT.ensure(colT);
  h->dzT.ensure(dzT);
  h->dzb.ensure(dzb);
  h->dA0.ensure(grad);
  h->dA1.ensure(grad);
  h->dcol.ensure(dcol);
  h->wb.ensure(wmax);
  h->wbT.ensure(wmax);
  const long long cells = (long long)batch * h->G * h->G;
  h->dlog.ensure((size_t)cells * round16(h->Nh));
  h->cell_loss.ensure((size_t)cells * 2);
  h->heat_t.ensure((size_t)cells);
  h->bin_t.ensure((size_t)cells);
}

// New parameters (host) with zero Adam moments, on the handle's stream: the enqueued work that still
// reads or writes them (an Adam step does not wait) finishes first, and the copies complete before
// the call returns.
void reset_params(tm_cnn* h, const float* params) {
  CNN_HIP(hipStreamSynchronize(h->stream));
  CNN_HIP(hipMemcpyAsync(h->w.p, params, h->n_params * sizeof(float), hipMemcpyHostToDevice, h->stream));
  CNN_HIP(hipMemsetAsync(h->m.p, 0, h->n_params * sizeof(float), h->stream));
  CNN_HIP(hipMemsetAsync(h->v.p, 0, h->n_params * sizeof(float), h->stream));
  CNN_HIP(hipStreamSynchronize(h->stream));
  h->adam_t = 0;
}

// defined after the entry points that use them
void read_loss(tm_cnn* h, float* loss);
void forward_impl(tm_cnn* h, const float* src, int batch);
void backward_impl(tm_cnn* h, const float* heat_t, const int* bin_t, long long n_bin, float pos_weight);

}  // namespace

extern "C" {

const char* tm_cnn_last_error(void) { return g_cerr.c_str(); }

int tm_cnn_create(int device_ordinal, const tm_cnn_config* c, tm_cnn** out) {
  return cguard([&] {
    CNN_REQUIRE(c && out, TM_ERR_INVALID, "NULL argument");
    *out = nullptr;
    CNN_REQUIRE(c->n_layers >= 1 && c->n_layers <= TM_CNN_MAX_LAYERS, TM_ERR_INVALID, "n_layers must be 1..16");
    CNN_REQUIRE(c->n_bins >= 1 && c->max_batch >= 1, TM_ERR_INVALID, "n_bins and max_batch must be positive");
    CNN_REQUIRE(c->input_size >= 4, TM_ERR_INVALID, "input_size too small");
    int dev_count = 0;
    CNN_HIP(hipGetDeviceCount(&dev_count));
    CNN_REQUIRE(device_ordinal >= 0 && device_ordinal < dev_count, TM_ERR_INVALID, "no such device");
    auto h = std::make_unique<tm_cnn>();
    h->device = device_ordinal;
    h->cfg = *c;
    CNN_HIP(hipSetDevice(device_ordinal));
    int s = c->input_size, cin = 1;
    long long off = 0;
    for (int l = 0; l < c->n_layers; ++l) {
      CNN_REQUIRE(c->strides[l] == 1 || c->strides[l] == 2, TM_ERR_INVALID, "strides must be 1 or 2");
      CNN_REQUIRE(c->channels[l] >= 1 && c->channels[l] <= 4096, TM_ERR_INVALID, "channels must be 1..4096");
      CNN_REQUIRE(s % c->strides[l] == 0, TM_ERR_INVALID, "input_size not divisible by the strides");
      Layer L;
      L.cin = cin;
      L.cout = c->channels[l];
      L.stride = c->strides[l];
      L.si = s;
      L.so = s / L.stride;
      L.K = 9 * cin;
      L.Kp = round16(L.K);
      L.w_off = off;
      off += (long long)L.cout * L.K;
      L.b_off = off;
      off += L.cout;
      h->layers.push_back(L);
      s = L.so;
      cin = L.cout;
    }
    h->G = s;
    h->Nh = 1 + c->n_bins;
    Layer hd;
    hd.cin = cin;
    hd.cout = h->Nh;
    hd.stride = 1;
    hd.si = hd.so = s;
    hd.K = cin;
    hd.Kp = round16(cin);
    hd.w_off = off;
    off += (long long)hd.cout * hd.K;
    hd.b_off = off;
    off += hd.cout;
    h->layers.push_back(hd);
    h->n_params = off;
    h->act.resize(h->layers.size() + 1);
    CNN_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    h->w.ensure(off);
    h->g.ensure(off);
    h->m.ensure(off);
    h->v.ensure(off);
    CNN_HIP(hipMemsetAsync(h->g.p, 0, off * sizeof(float), h->stream));
    alloc_for_batch(h.get(), c->max_batch);
    h->batch = 0;
    CNN_HIP(hipStreamSynchronize(h->stream));
    *out = h.release();
  });
}

int tm_cnn_destroy(tm_cnn* h) {
  return cguard([&] {
    if (!h) return;
    (void)hipSetDevice(h->device);
    delete h;  // the destructor waits for and destroys the stream, the buffers free themselves
  });
}

int tm_cnn_shape(tm_cnn* h, long long* n_params, int* grid) {
  return cguard([&] {
    CNN_REQUIRE(h, TM_ERR_INVALID, "handle is NULL");
    if (n_params) *n_params = h->n_params;
    if (grid) *grid = h->G;
  });
}

int tm_cnn_set_params(tm_cnn* h, const float* params) {
  return cguard([&] {
    require(h);
    CNN_REQUIRE(params, TM_ERR_INVALID, "params is NULL");
    reset_params(h, params);
  });
}

int tm_cnn_get_params(tm_cnn* h, float* params) {
  return cguard([&] {
    require(h);
    CNN_REQUIRE(params, TM_ERR_INVALID, "params is NULL");
    CNN_HIP(hipStreamSynchronize(h->stream));
    CNN_HIP(hipMemcpy(params, h->w.p, h->n_params * sizeof(float), hipMemcpyDeviceToHost));
  });
}

int tm_cnn_init_params(tm_cnn* h, unsigned long long seed) {
  return cguard([&] {
    require(h);
    std::vector<float> p((size_t)h->n_params, 0.f);
    std::mt19937_64 rng(seed);
    for (const Layer& L : h->layers) {
      std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / (double)L.K));
      for (long long e = 0; e < (long long)L.cout * L.K; ++e) p[L.w_off + e] = (float)nd(rng);
    }
    reset_params(h, p.data());
  });
}

int tm_cnn_forward(tm_cnn* h, const float* images, int batch, float* logits) {
  return cguard([&] {
    require(h);
    CNN_REQUIRE(images, TM_ERR_INVALID, "images is NULL");
    CNN_REQUIRE(batch >= 1 && batch <= h->cfg.max_batch, TM_ERR_INVALID, "batch must be 1..max_batch");
    const long long npx = (long long)h->cfg.input_size * h->cfg.input_size;
    CNN_HIP(hipMemcpyAsync(h->act[0].p, images, batch * npx * sizeof(float), hipMemcpyHostToDevice, h->stream));
    forward_impl(h, h->act[0].p, batch);
    if (logits) {
      const long long cells = (long long)batch * h->G * h->G;
      CNN_HIP(hipMemcpyAsync(logits, h->act.back().p, cells * h->Nh * sizeof(float), hipMemcpyDeviceToHost,
                             h->stream));
    }
    CNN_HIP(hipStreamSynchronize(h->stream));
  });
}

int tm_cnn_forward_device(tm_cnn* h, const float* images, int batch) {
  return cguard([&] {
    require(h);
    CNN_REQUIRE(images, TM_ERR_INVALID, "images is NULL");
    CNN_REQUIRE(batch >= 1 && batch <= h->cfg.max_batch, TM_ERR_INVALID, "batch must be 1..max_batch");
    forward_impl(h, images, batch);
  });
}

int tm_cnn_backward(tm_cnn* h, const float* heat_target, const int* bin_target, float pos_weight, float* loss,
                    float* grads) {
  return cguard([&] {
    require(h);
    CNN_REQUIRE(heat_target && bin_target, TM_ERR_INVALID, "NULL target");
    CNN_REQUIRE(h->batch > 0, TM_ERR_STATE, "no forward pass to differentiate");
    hipStream_t st = h->stream;
    const long long cells = (long long)h->batch * h->G * h->G;
    long long n_bin = 0;
    for (long long c = 0; c < cells; ++c) {
      CNN_REQUIRE(bin_target[c] < h->cfg.n_bins, TM_ERR_INVALID, "bin target out of range");
      n_bin += bin_target[c] >= 0;
    }
    CNN_HIP(hipMemcpyAsync(h->heat_t.p, heat_target, cells * sizeof(float), hipMemcpyHostToDevice, st));
    CNN_HIP(hipMemcpyAsync(h->bin_t.p, bin_target, cells * sizeof(int), hipMemcpyHostToDevice, st));
    backward_impl(h, h->heat_t.p, h->bin_t.p, n_bin, pos_weight);
    if (grads) CNN_HIP(hipMemcpyAsync(grads, h->g.p, h->n_params * sizeof(float), hipMemcpyDeviceToHost, st));
    if (loss) read_loss(h, loss);
    CNN_HIP(hipStreamSynchronize(st));
  });
}

int tm_cnn_backward_device(tm_cnn* h, const float* heat_target, const int* bin_target, long long n_labelled,
                           float pos_weight) {
  return cguard([&] {
    require(h);
    CNN_REQUIRE(heat_target && bin_target, TM_ERR_INVALID, "NULL target");
    CNN_REQUIRE(h->batch > 0, TM_ERR_STATE, "no forward pass to differentiate");
    CNN_REQUIRE(n_labelled >= 0, TM_ERR_INVALID, "n_labelled must be >= 0");
    backward_impl(h, heat_target, bin_target, n_labelled, pos_weight);
  });
}

int tm_cnn_loss(tm_cnn* h, float* loss) {
  return cguard([&] {
    require(h);
    CNN_REQUIRE(loss, TM_ERR_INVALID, "loss is NULL");
    CNN_REQUIRE(h->loss_cells > 0, TM_ERR_STATE, "no backward pass");
    read_loss(h, loss);
  });
}

int tm_cnn_synchronize(tm_cnn* h) {
  return cguard([&] {
    require(h);
    CNN_HIP(hipStreamSynchronize(h->stream));
  });
}
}  // extern "C"

namespace {
// the two loss terms of the last backward pass (waits for it): means of the per-cell terms over all
// cells and over the cells with a bin
void read_loss(tm_cnn* h, float* loss) {
  const long long cells = h->loss_cells;
  h->h_loss.resize((size_t)cells * 2);
  CNN_HIP(hipMemcpyAsync(h->h_loss.data(), h->cell_loss.p, cells * 2 * sizeof(float), hipMemcpyDeviceToHost,
                         h->stream));
  CNN_HIP(hipStreamSynchronize(h->stream));
  double lh = 0.0, lb = 0.0;
  for (long long c = 0; c < cells; ++c) {
    lh += h->h_loss[2 * c];
    lb += h->h_loss[2 * c + 1];
  }
  loss[0] = (float)(lh / (double)cells);
  loss[1] = (float)(lb / (double)std::max(1LL, h->loss_bins));
}

// forward from the raw images at src (device memory; act[0] itself for the host entry point)
void forward_impl(tm_cnn* h, const float* src, int batch) {
  h->batch = batch;
  const long long npx = (long long)h->cfg.input_size * h->cfg.input_size;
  h->std_part.ensure((size_t)batch * STD_PARTS * 2);
  hipLaunchKernelGGL(k_std_part, dim3(STD_PARTS, batch), dim3(256), 0, h->stream, src, npx, h->std_part.p);
  check_launch("k_std_part");
  hipLaunchKernelGGL(k_std_apply, dim3((unsigned)std::min<long long>((npx + 1023) / 1024, 1024), batch), dim3(256), 0,
                     h->stream, src, h->act[0].p, npx, h->std_part.p);
  check_launch("k_std_apply");
  for (int l = 0; l < (int)h->layers.size(); ++l) forward_layer(h, l);
}

// loss and gradients of the last forward pass from device targets; n_bin cells carry a bin
void backward_impl(tm_cnn* h, const float* heat_t, const int* bin_t, long long n_bin, float pos_weight) {
  hipStream_t st = h->stream;
  const int B = h->batch, Nh = h->Nh;
  const long long cells = (long long)B * h->G * h->G;
  h->loss_cells = cells;
  h->loss_bins = n_bin;
  const float heat_scale = 1.f / (float)cells, bin_scale = 1.f / (float)std::max(1LL, n_bin);
  hipLaunchKernelGGL(k_loss, dim3((unsigned)cells), dim3(64), 0, st, h->act.back().p, Nh, heat_t, bin_t,
                     pos_weight, heat_scale, bin_scale, h->dlog.p, round16(Nh), h->cell_loss.p);
  check_launch("k_loss");
  // walk the layers backwards; dcur holds dL/d(output of layer l) (pre-activation for the head)
  float* dcur = h->dlog.p;
  float* dnext = h->dA0.p;
  bool masked = false;  // dcur already carries the layer's ReLU backward (k_conv_dgrad's epilogue)
  for (int l = (int)h->layers.size() - 1; l >= 0; --l) {
    const Layer& L = h->layers[l];
    const bool head = l == (int)h->layers.size() - 1;
    const long long M = rows_of(h, L), Mp = round16ll(M);
    if (!head && !masked) {
      hipLaunchKernelGGL(k_relu_bwd, dim3(grid1(M * L.cout)), dim3(256), 0, st, dcur, h->act[l + 1].p, M * L.cout);
      check_launch("k_relu_bwd");
    }
    // dcur rows: the head's gradient dL/dlogits is padded to a multiple of 16 columns (zeros)
    const int ldz = head ? round16(L.cout) : L.cout;
    if (wgrad_implicit(L, B, ldz, head)) {
      // weights and biases together (the biases follow the weights in the flat gradient)
      if (head)
        conv_wgrad<1>(st, h->parts, h->act[l].p, L, B, dcur, ldz, h->g.p + L.w_off);
      else if (L.cin == 1)
        conv_wgrad<0>(st, h->parts, h->act[l].p, L, B, dcur, ldz, h->g.p + L.w_off);
      else
        conv_wgrad<9>(st, h->parts, h->act[l].p, L, B, dcur, ldz, h->g.p + L.w_off);
    } else {
      colsum(st, h->csum, dcur, M, L.cout, ldz, h->g.p + L.b_off);
      // the layer's input as the product operand (recomputed), then its transpose [Kp][Mp]
      if (head)
        to_bf16_pad(st, h->act[l].p, M, L.cin, L.cin, h->col.p, L.Kp);
      else
        im2col(st, h->act[l].p, L, M, h->col.p);
      transpose_bf16(st, h->col.p, M, L.Kp, L.Kp, h->colT.p, Mp, L.Kp, Mp);
      // dZ^T [cout][Mp] and dW[cout][K] = dZ^T . col^T^T
      transpose_bf16(st, dcur, M, L.cout, ldz, h->dzT.p, Mp, L.cout, Mp);
      gemm(st, h->parts, h->dzT.p, (int)Mp, h->colT.p, (int)Mp, h->g.p + L.w_off, L.K, L.cout, L.K, (int)Mp,
           nullptr, false);
    }
    if (l == 0) break;
    if (!head && L.cout % 16 == 0) {
      // implicit transposed convolution straight into the input's gradient
      pack_b<PACK_DGRAD>(st, h->w.p + L.w_off, L.cin, 9 * L.cout, 0, 0, L.cin, L.cout, h->wbT.p);
      conv_dgrad(st, dcur, L, B, h->wbT.p, h->act[l].p, dnext);
      masked = true;
      dcur = dnext;
      dnext = (dnext == h->dA0.p) ? h->dA1.p : h->dA0.p;
      continue;
    }
    // dcol[M][K] = dZ[M][cout16] . W^T[K][cout16]^T, then the gather into the input's gradient
    const int c16 = round16(L.cout);
    if (head && L.cin % 16 == 0) {
      // the head's data gradient as a 1x1 implicit product over its padded dL/dlogits rows (B = W^T,
      // zero past the head's outputs), with the last convolution's ReLU backward in the epilogue
      pack_b<PACK_COLS>(st, h->w.p + L.w_off, L.cin, ldz, L.cin, L.cout, 0, 0, h->wbT.p);
      conv_fwd<1, EPI_MASK>(st, dcur, L.si, ldz, L.so, 1, h->wbT.p, dnext, M, L.cin, nullptr, h->act[l].p);
      masked = true;
      dcur = dnext;
      dnext = (dnext == h->dA0.p) ? h->dA1.p : h->dA0.p;
      continue;
    }
    transpose_bf16(st, h->w.p + L.w_off, L.cout, L.K, L.K, h->wbT.p, c16, L.K, c16);
    to_bf16_pad(st, dcur, M, L.cout, ldz, h->dzb.p, c16);
    if (head) {
      gemm(st, h->parts, h->dzb.p, c16, h->wbT.p, c16, dnext, L.K, M, L.K, c16, nullptr, false);
    } else {
      gemm(st, h->parts, h->dzb.p, c16, h->wbT.p, c16, h->dcol.p, L.K, M, L.K, c16, nullptr, false);
      const long long total = (long long)B * L.si * L.si * L.cin;
      CNN_REQUIRE(total < (1LL << 31), TM_ERR_INVALID, "col2im: layer too large for one launch");
      const long long threads = (L.cin % 4 == 0) ? total / 4 : total;
      hipLaunchKernelGGL(k_col2im, dim3(grid1(threads)), dim3(256), 0, st, h->dcol.p, L.K, L.si, L.cin, L.so,
                         L.stride, (int)total, dnext);
      check_launch("k_col2im");
    }
    masked = false;
    dcur = dnext;
    dnext = (dnext == h->dA0.p) ? h->dA1.p : h->dA0.p;
  }
}
}  // namespace

extern "C" {

int tm_cnn_heat(tm_cnn* h, float* heat) {
  return cguard([&] {
    require(h);
    CNN_REQUIRE(heat, TM_ERR_INVALID, "heat is NULL");
    CNN_REQUIRE(h->batch > 0, TM_ERR_STATE, "no forward pass");
    const long long cells = (long long)h->batch * h->G * h->G;
    // column 0 of the cell-major logits, gathered on the device and copied in one piece
    h->gathered.ensure((size_t)cells);
    hipLaunchKernelGGL(k_column0, dim3(grid1(cells)), dim3(256), 0, h->stream, h->act.back().p, h->Nh, cells,
                       h->gathered.p);
    check_launch("k_column0");
    CNN_HIP(hipMemcpyAsync(heat, h->gathered.p, cells * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    CNN_HIP(hipStreamSynchronize(h->stream));
  });
}

int tm_cnn_cell_logits(tm_cnn* h, const int* cells, int n, float* out) {
  return cguard([&] {
    require(h);
    CNN_REQUIRE(h->batch > 0, TM_ERR_STATE, "no forward pass");
    CNN_REQUIRE(n >= 0 && (n == 0 || (cells && out)), TM_ERR_INVALID, "NULL argument");
    if (n == 0) return;
    const long long total = (long long)h->batch * h->G * h->G;
    for (int q = 0; q < n; ++q) CNN_REQUIRE(cells[q] >= 0 && cells[q] < total, TM_ERR_INVALID, "cell out of range");
    h->cells.ensure(n);
    h->gathered.ensure((size_t)n * h->Nh);
    CNN_HIP(hipMemcpyAsync(h->cells.p, cells, n * sizeof(int), hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_gather_cells, dim3(n), dim3(256), 0, h->stream, h->act.back().p, h->Nh, h->cells.p, n,
                       h->gathered.p);
    check_launch("k_gather_cells");
    CNN_HIP(hipMemcpyAsync(out, h->gathered.p, (size_t)n * h->Nh * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    CNN_HIP(hipStreamSynchronize(h->stream));
  });
}

int tm_cnn_adam_step(tm_cnn* h, float lr, float beta1, float beta2, float eps) {
  return cguard([&] {
    require(h);
    CNN_REQUIRE(lr > 0.f && beta1 >= 0.f && beta1 < 1.f && beta2 >= 0.f && beta2 < 1.f && eps > 0.f, TM_ERR_INVALID,
                "invalid Adam hyper-parameters");
    h->adam_t += 1;
    const double c1 = 1.0 - std::pow((double)beta1, (double)h->adam_t);
    const double c2 = 1.0 - std::pow((double)beta2, (double)h->adam_t);
    const float lr_t = (float)(lr * std::sqrt(c2) / c1);
    hipLaunchKernelGGL(k_adam, dim3(grid1(h->n_params)), dim3(256), 0, h->stream, h->w.p, h->g.p, h->m.p, h->v.p,
                       h->n_params, lr_t, beta1, beta2, eps);
    check_launch("k_adam");
  });
}

int tm_cnn_get_activation(tm_cnn* h, int layer, float* out) {
  return cguard([&] {
    require(h);
    CNN_REQUIRE(out, TM_ERR_INVALID, "out is NULL");
    CNN_REQUIRE(h->batch > 0, TM_ERR_STATE, "no forward pass");
    CNN_REQUIRE(layer >= 0 && layer < (int)h->layers.size(), TM_ERR_INVALID, "layer out of range");
    const Layer& L = h->layers[layer];
    const size_t n = (size_t)rows_of(h, L) * L.cout;
    CNN_HIP(hipStreamSynchronize(h->stream));
    CNN_HIP(hipMemcpy(out, h->act[layer + 1].p, n * sizeof(float), hipMemcpyDeviceToHost));
  });
}

int tm_cnn_debug_gemm(int device_ordinal, const float* A, const float* B, int M, int N, int K, float* C) {
  return cguard([&] {
    CNN_REQUIRE(A && B && C && M > 0 && N > 0 && K > 0, TM_ERR_INVALID, "invalid arguments");
    CNN_HIP(hipSetDevice(device_ordinal));
    hipStream_t st;
    CNN_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct S {
      hipStream_t s;
      ~S() { (void)hipStreamDestroy(s); }
    } guard{st};
    const int Kp = round16(K);
    Buf<float> a, b, c, parts;
    Buf<bf16_t> ab, bb;
    a.ensure((size_t)M * K);
    b.ensure((size_t)N * K);
    c.ensure((size_t)M * N);
    ab.ensure((size_t)M * Kp);
    bb.ensure((size_t)N * Kp);
    CNN_HIP(hipMemcpyAsync(a.p, A, (size_t)M * K * sizeof(float), hipMemcpyHostToDevice, st));
    CNN_HIP(hipMemcpyAsync(b.p, B, (size_t)N * K * sizeof(float), hipMemcpyHostToDevice, st));
    to_bf16_pad(st, a.p, M, K, K, ab.p, Kp);
    to_bf16_pad(st, b.p, N, K, K, bb.p, Kp);
    gemm(st, parts, ab.p, Kp, bb.p, Kp, c.p, N, M, N, Kp, nullptr, false);
    CNN_HIP(hipMemcpyAsync(C, c.p, (size_t)M * N * sizeof(float), hipMemcpyDeviceToHost, st));
    CNN_HIP(hipStreamSynchronize(st));
  });
}

}  // extern "C"
