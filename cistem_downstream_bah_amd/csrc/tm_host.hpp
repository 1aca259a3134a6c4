// Host-side pieces shared by the search (tm_core.hip) and the refinement (tm_refine.hip): error
// plumbing, twiddle and quadrant-swap tables, the Euler matrix, device buffers and the
// size-dispatched launches of the FFT kernels. A launch context is any type with the members
// `hipStream_t stream` and `const cf* twiddle(int n)` (a cached device twiddle table).
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tm_hip.h"
#include "tm_kernels.hpp"
#include "tm_sizes.hpp"

using tmk::cf;

namespace {

thread_local std::string g_err;

struct TmError {
  int code;
  std::string msg;
};

#define TM_HIP(call)                                                                                   \
  do {                                                                                                 \
    hipError_t e_ = (call);                                                                            \
    if (e_ != hipSuccess)                                                                              \
      throw TmError{TM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)};                  \
  } while (0)

#define TM_REQUIRE(cond, code, msg)                                                                    \
  do {                                                                                                 \
    if (!(cond)) throw TmError{code, msg};                                                             \
  } while (0)

template <class F>
int guarded(F&& f) {
  try {
    f();
    return TM_OK;
  } catch (const TmError& e) {
    g_err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return TM_ERR_INVALID;
  }
}

// Accurate twiddle table exp(-2 pi i k / n), k < n.
std::vector<cf> twiddles(int n) {
  std::vector<cf> t(n);
  const long double two_pi = 2.0L * acosl(-1.0L);
  for (int k = 0; k < n; ++k) {
    long double a = two_pi * (long double)k / (long double)n;
    t[k].x = (float)cosl(a);
    t[k].y = (float)-sinl(a);
  }
  return t;
}

// Quadrant-swap phase factors of Image::SwapRealSpaceQuadrants on an n x n image that is not
// centred (shift n/2 per axis, image.cpp:9757-9822 -> PhaseShift image.cpp:9221-9268), with
// the reference's float phase arithmetic (functions.h:494-504).
std::vector<cf> swap_phase(int n) {
  const double PI = 3.14159265359;
  const int nh = n / 2 + 1;
  std::vector<cf> t((size_t)n * nh);
  const float shift = (float)(n / 2);
  for (int j = 0; j < n; ++j) {
    const int ly = j >= n / 2 ? j - n : j;
    const float py = (float)(shift * (float)ly * 2.0 * PI / (float)n);
    for (int i = 0; i < nh; ++i) {
      const float px = (float)(shift * (float)i * 2.0 * PI / (float)n);
      const float pz = 0.0f;
      const float ph = -px - py - pz;
      t[(size_t)j * nh + i].x = cosf(ph);
      t[(size_t)j * nh + i].y = sinf(ph);
    }
  }
  return t;
}

// deg_2_rad and the ZYZ Euler matrix of RotationMatrix::SetToEulerRotation (matrix.cpp:490-515),
// row-major m[3*r + c].
void euler_matrix(float phi, float theta, float psi, float* m) {
  const float PIf = 3.14159265359f;
  auto d2r = [&](float d) { return (float)(d * PIf / 180.); };
  const float cp = cosf(d2r(phi)), sp = sinf(d2r(phi));
  const float ct = cosf(d2r(theta)), st = sinf(d2r(theta));
  const float cs = cosf(d2r(psi)), ss = sinf(d2r(psi));
  m[0] = cp * ct * cs - sp * ss;
  m[3] = sp * ct * cs + cp * ss;
  m[6] = -st * cs;
  m[1] = -cp * ct * ss - sp * cs;
  m[4] = -sp * ct * ss + cp * cs;
  m[7] = st * ss;
  m[2] = st * cp;
  m[5] = st * sp;
  m[8] = ct;
}

// A limit on the device bytes the calling thread may allocate while it is active (tm_init's
// device_memory_limit); exceeding it fails like hipMalloc running out of memory.
struct AllocLimit {
  long long limit = 0, used = 0;
};
thread_local AllocLimit* g_alloc_limit = nullptr;

inline void device_alloc(void** p, size_t bytes) {
  if (g_alloc_limit) {
    if (g_alloc_limit->used + (long long)bytes > g_alloc_limit->limit)
      throw TmError{TM_ERR_HIP, "device allocation of " + std::to_string(bytes) + " bytes would exceed the limit of " +
                                    std::to_string(g_alloc_limit->limit) + " bytes (" +
                                    std::to_string(g_alloc_limit->used) + " allocated): out of memory"};
    g_alloc_limit->used += (long long)bytes;
  }
  TM_HIP(hipMalloc(p, bytes));
}

// Activates a limit for the current scope (no limit when bytes <= 0).
struct ScopedAllocLimit {
  AllocLimit lim;
  AllocLimit* prev;
  explicit ScopedAllocLimit(long long bytes) : prev(g_alloc_limit) {
    lim.limit = bytes;
    if (bytes > 0) g_alloc_limit = &lim;
  }
  ~ScopedAllocLimit() { g_alloc_limit = prev; }
};

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    device_alloc((void**)&p, count * sizeof(T));
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
};

// ------------------------------------------------------------------ size dispatch
void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw TmError{TM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e)};
}

// ------------------------------------------------------------------ size dispatch
template <int N, class Ctx>
void launch_cols(Ctx* h, cf* data, int ncols, size_t bstride, int nb, int sign, int mode, const cf* filt,
                 float scale) {
  constexpr int NT = 64;
  const cf* tw = h->twiddle(N);
  const dim3 g(ncols * nb);
  if (mode == 1)
    hipLaunchKernelGGL((tmk::k_cols<N, NT, -1, 1>), g, dim3(NT), 0, h->stream, data, ncols, bstride, tw, filt, scale);
  else if (sign < 0)
    hipLaunchKernelGGL((tmk::k_cols<N, NT, -1, 0>), g, dim3(NT), 0, h->stream, data, ncols, bstride, tw, filt, scale);
  else
    hipLaunchKernelGGL((tmk::k_cols<N, NT, +1, 0>), g, dim3(NT), 0, h->stream, data, ncols, bstride, tw, filt, scale);
  check_launch("k_cols");
}

template <class Ctx>
void cols(Ctx* h, int N, cf* data, int ncols, size_t bstride, int nb, int sign, int mode, const cf* filt,
          float scale) {
  switch (N) {
#define TM_CASE(n) \
  case n: launch_cols<n, Ctx>(h, data, ncols, bstride, nb, sign, mode, filt, scale); break;
    TM_TEMPLATE_SIZES(TM_CASE)
#undef TM_CASE
    default: throw TmError{TM_ERR_UNSUPPORTED, "unsupported column FFT length " + std::to_string(N)};
  }
}

template <int N>
constexpr int row_threads() { return (N / 2) >= 1024 ? 256 : 64; }

template <int N, class Ctx>
void launch_c2r(Ctx* h, const cf* in, size_t ibs, int irs, float* out, size_t obs, int ors, int nrows, int nb) {
  constexpr int NT = row_threads<N>();
  hipLaunchKernelGGL((tmk::k_rows_c2r<N, NT>), dim3(nrows, nb), dim3(NT), 0, h->stream, in, ibs, irs, out, obs, ors,
                     h->twiddle(N), h->twiddle(N / 2));
  check_launch("k_rows_c2r");
}

template <class Ctx>
void rows_c2r(Ctx* h, int N, const cf* in, size_t ibs, int irs, float* out, size_t obs, int ors, int nrows,
              int nb) {
  switch (N) {
#define TM_CASE(n) \
  case n: launch_c2r<n, Ctx>(h, in, ibs, irs, out, obs, ors, nrows, nb); break;
    TM_TEMPLATE_SIZES(TM_CASE)
#undef TM_CASE
    default: throw TmError{TM_ERR_UNSUPPORTED, "unsupported C2R row length " + std::to_string(N)};
  }
}

// C2R rows of any compiled line length (the template sizes and the image sides).
template <class Ctx>
void rows_c2r_line(Ctx* h, int N, const cf* in, size_t ibs, int irs, float* out, size_t obs, int ors, int nrows,
                   int nb) {
  switch (N) {
#define TM_CASE(n) \
  case n: launch_c2r<n, Ctx>(h, in, ibs, irs, out, obs, ors, nrows, nb); break;
    TM_LINE_SIZES(TM_CASE)
#undef TM_CASE
    default: throw TmError{TM_ERR_UNSUPPORTED, "unsupported C2R row length " + std::to_string(N)};
  }
}

template <int N, class Ctx, int OS = tmk::ZS_F32>
void launch_r2c(Ctx* h, const float* src, size_t sbs, int srs, int r0, int s0, int len, int c0, cf* dst,
                size_t dbs, int drs, int nrows, int nb, float scale) {
  constexpr int NT = row_threads<N>();
  hipLaunchKernelGGL((tmk::k_rows_r2c<N, NT, OS>), dim3(nrows, nb), dim3(NT), 0, h->stream, src, sbs, srs, r0, s0, len,
                     c0, dst, dbs, drs, h->twiddle(N), h->twiddle(N / 2), scale);
  check_launch("k_rows_r2c");
}

// Row R2C of an image width with 16-bit output (OS = tmk::ZS_F16 or ZS_BF16): the search's forward
// rows when its plane is stored in 16 bits.
template <int OS, class Ctx>
void rows_r2c_16(Ctx* h, int N, const float* src, size_t sbs, int srs, int r0, int s0, int len, int c0, cf* dst,
                 size_t dbs, int drs, int nrows, int nb, float scale) {
  switch (N) {
#define TM_CASE(n) \
  case n: launch_r2c<n, Ctx, OS>(h, src, sbs, srs, r0, s0, len, c0, dst, dbs, drs, nrows, nb, scale); break;
    TM_IMAGE_SIZES(TM_CASE)
#undef TM_CASE
    default: throw TmError{TM_ERR_UNSUPPORTED, "unsupported R2C row length " + std::to_string(N)};
  }
}

template <class Ctx>
void rows_r2c(Ctx* h, int N, const float* src, size_t sbs, int srs, int r0, int s0, int len, int c0, cf* dst,
              size_t dbs, int drs, int nrows, int nb, float scale) {
  switch (N) {
#define TM_CASE(n) \
  case n: launch_r2c<n, Ctx>(h, src, sbs, srs, r0, s0, len, c0, dst, dbs, drs, nrows, nb, scale); break;
    TM_LINE_SIZES(TM_CASE)
#undef TM_CASE
    default: throw TmError{TM_ERR_UNSUPPORTED, "unsupported R2C row length " + std::to_string(N)};
  }
}

template <int N, int LAYOUT, class Ctx>
void launch_extract_cols(Ctx* h, const cf* tpl, const float* mats, int o0, int nb, const cf* phase, const cf* filt,
                         cf* out) {
  constexpr int KC = 8;
  const dim3 g((N / 2 + 1 + KC - 1) / KC, nb);
  hipLaunchKernelGGL((tmk::k_extract_cols<N, KC, LAYOUT>), g, dim3(KC * 64), 0, h->stream, tpl, mats, o0, phase, filt,
                     h->twiddle(N), out);
  check_launch("k_extract_cols");
}

// Slice extraction and the inverse column FFT of the slice (P != T) or of the filtered projection
// (P == T), fused. LAYOUT: the template spectrum's layout (tpl_index in tm_kernels.hpp).
template <int LAYOUT = 0, class Ctx>
void extract_cols(Ctx* h, int N, const cf* tpl, const float* mats, int o0, int nb, const cf* phase, const cf* filt,
                  cf* out) {
  switch (N) {
#define TM_CASE(n) \
  case n: launch_extract_cols<n, LAYOUT, Ctx>(h, tpl, mats, o0, nb, phase, filt, out); break;
    TM_TEMPLATE_SIZES(TM_CASE)
#undef TM_CASE
    default: throw TmError{TM_ERR_UNSUPPORTED, "unsupported slice size " + std::to_string(N)};
  }
}

// Column transforms of any compiled line length (the template sizes and the image sides).
template <class Ctx>
void cols_line(Ctx* h, int N, cf* data, int ncols, size_t bstride, int nb, int sign) {
  switch (N) {
#define TM_CASE(n) \
  case n: launch_cols<n, Ctx>(h, data, ncols, bstride, nb, sign, 0, nullptr, 1.0f); break;
    TM_LINE_SIZES(TM_CASE)
#undef TM_CASE
    default: throw TmError{TM_ERR_UNSUPPORTED, "unsupported column FFT length " + std::to_string(N)};
  }
}

}  // namespace
