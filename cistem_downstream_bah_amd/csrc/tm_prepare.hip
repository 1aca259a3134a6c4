// Caller-side preparation of the search inputs (the host steps of match_template.cpp that build
// the arguments of TemplateMatchingCore::Init / RunInnerLoop), with the large transforms on the GPU:
//   tm_prepare_sizes / tm_prepare_input_image   match_template.cpp:490-598, 703-724
//   tm_prepare_template                         match_template.cpp:621-625, 837-845
//   tm_prepare_template_scaled                  match_template.cpp:837-845 with ChangePixelSize != 1
//                                               (image.cpp:8705-8792)
//   tm_projection_filter                        match_template.cpp:893-898 (CTF x whitening)
//   tm_prepare_refine_image                     refine_template.cpp:526-546 (whitened real-space image)
//   tm_search_grid / tm_psi_list                euler_search.cpp:180-282, 353-443; match_template.cpp:658-697, 1001
// Host loops restate the reference's Image/Curve/CTF arithmetic; the 2-D and 3-D R2C transforms
// run on the device with the same kernels as the search.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tm_hip.h"
#include "../../include/tm_prepare.h"
#include "tm_kernels.hpp"
#include "tm_sizes.hpp"

using tmk::cf;

namespace {

thread_local std::string g_perr;
struct PErr {
  int code;
  std::string msg;
};
#define P_HIP(call)                                                                          \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) throw PErr{TM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
  } while (0)
#define P_REQ(c, code, m) \
  do {                    \
    if (!(c)) throw PErr{code, m}; \
  } while (0)

template <class F>
int pguard(F&& f) {
  try {
    f();
    return TM_OK;
  } catch (const PErr& e) {
    g_perr = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    g_perr = e.what();
    return TM_ERR_INVALID;
  }
}

const double PI = 3.14159265359;   // defines.h:6
const float PIf = 3.14159265359f;  // defines.h:7
inline float deg_2_rad(float d) { return d * PIf / 180.; }
inline float rad_2_deg(float r) { return r / (PIf / 180.); }
inline bool even(int n) { return (n & 1) == 0; }
inline bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
inline bool line_size_compiled(int n) {
  switch (n) {
#define TM_CASE(k) case k:
    TM_LINE_SIZES(TM_CASE)
#undef TM_CASE
    return true;
    default: return false;
  }
}
inline bool is_power_of_two(int n) {  // functions.h:422-426
  if (n == 0) return false;
  return std::ceil(std::log2((float)n)) == std::floor(std::log2((float)n));
}

// ---------------------------------------------------------------- functions.cpp:529-627
int factorized(int wanted, int largest_factor, int step, bool upper) {
  int t = wanted;
  if (!even(t)) t += upper ? 1 : -1;
  for (int number = t; upper ? number < 10000 * wanted : number >= 2; number += upper ? step : -step) {
    int rem = number;
    for (int f = 2; f <= largest_factor && rem != 1; ++f)
      while (rem % f == 0) rem /= f;
    if (rem == 1) return number;
  }
  return upper ? 10000 * wanted : 0;
}

// match_template.cpp:527-560 (enforce_factor = (int)true = 1, so the search step is 1)
int factorize_axis(int original, int reference_size) {
  const int primes[6] = {2, 3, 5, 7, 9, 13};
  for (int i = 0; i < 6; ++i) {
    const int neg = factorized(original, primes[i], 1, false);
    const int pos = factorized(original, primes[i], 1, true);
    if ((float)(original - neg) < (float)reference_size * 0.000001f) return neg;
    if ((float)(-original + pos) < (float)original * 0.1f) return pos;
  }
  return original;
}

// ---------------------------------------------------------------- real images in the padded layout
struct Img {
  int nx = 0, ny = 0, nz = 1;
  std::vector<float> v;  // rows of nx + 2 (even nx)
  int rs() const { return nx + 2; }
  void alloc(int x, int y, int z) {
    nx = x; ny = y; nz = z;
    v.assign((size_t)(x + 2) * y * z, 0.f);
  }
  float& at(int x, int y, int z = 0) { return v[((size_t)z * ny + y) * rs() + x]; }
  float at(int x, int y, int z = 0) const { return v[((size_t)z * ny + y) * rs() + x]; }
};

// Image::ReturnAverageOfRealValuesOnEdges (image.cpp:5907-6000)
float edge_mean(const Img& a) {
  double sum = 0.0;
  long n = 0;
  if (a.nz == 1) {
    for (int x = 0; x < a.nx; ++x) sum += a.at(x, 0);
    n += a.nx;
    for (int y = 1; y < a.ny - 1; ++y) { sum += a.at(0, y); sum += a.at(a.nx - 1, y); n += 2; }
    for (int x = 0; x < a.nx; ++x) sum += a.at(x, a.ny - 1);
    n += a.nx;
  } else {
    for (int y = 0; y < a.ny; ++y) for (int x = 0; x < a.nx; ++x) sum += a.at(x, y, 0);
    n += (long)a.nx * a.ny;
    for (int z = 1; z < a.nz - 1; ++z)
      for (int y = 0; y < a.ny; ++y) {
        if (y == 0 || y == a.ny - 1) { for (int x = 0; x < a.nx; ++x) sum += a.at(x, y, z); n += a.nx; }
        else { sum += a.at(0, y, z); sum += a.at(a.nx - 1, y, z); n += 2; }
      }
    for (int y = 0; y < a.ny; ++y) for (int x = 0; x < a.nx; ++x) sum += a.at(x, y, a.nz - 1);
    n += (long)a.nx * a.ny;
  }
  return sum / float(n);
}

// Image::ClipInto real-space branch (image.cpp:8522-8575), centres at dim/2
Img clip_into(const Img& a, int x, int y, int z, float pad) {
  Img o;
  o.alloc(x, y, z);
  for (int kk = 0; kk < z; ++kk) {
    const int k = a.nz / 2 + (kk - z / 2);
    for (int jj = 0; jj < y; ++jj) {
      const int j = a.ny / 2 + (jj - y / 2);
      for (int ii = 0; ii < x; ++ii) {
        const int i = a.nx / 2 + (ii - x / 2);
        o.at(ii, jj, kk) = (k < 0 || k >= a.nz || j < 0 || j >= a.ny || i < 0 || i >= a.nx) ? pad : a.at(i, j, k);
      }
    }
  }
  return o;
}

// ---------------------------------------------------------------- device transforms
struct Dev {
  hipStream_t s = nullptr;
  std::vector<void*> bufs;
  explicit Dev(int device) {
    P_HIP(hipSetDevice(device));
    P_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  }
  ~Dev() {
    if (s) (void)hipStreamSynchronize(s);
    for (void* p : bufs) (void)hipFree(p);
    if (s) (void)hipStreamDestroy(s);
  }
  template <class T>
  T* alloc(size_t n) {
    void* p = nullptr;
    P_HIP(hipMalloc(&p, n * sizeof(T)));
    bufs.push_back(p);
    return (T*)p;
  }
  cf* twiddles(int n) {
    std::vector<cf> t(n);
    const long double two_pi = 2.0L * acosl(-1.0L);
    for (int k = 0; k < n; ++k) {
      long double a = two_pi * (long double)k / (long double)n;
      t[k].x = (float)cosl(a);
      t[k].y = (float)-sinl(a);
    }
    cf* d = alloc<cf>(n);
    P_HIP(hipMemcpy(d, t.data(), n * sizeof(cf), hipMemcpyHostToDevice));
    return d;
  }
};

void chk(const char* w) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw PErr{TM_ERR_HIP, std::string(w) + ": " + hipGetErrorString(e)};
}

template <int N>
void rows_r2c_t(Dev& d, const float* src, size_t sbs, int srs, cf* dst, size_t dbs, int drs, int nrows, int nb) {
  constexpr int NT = (N / 2) >= 1024 ? 256 : 64;
  cf* twn = d.twiddles(N);
  cf* twh = d.twiddles(N / 2);
  hipLaunchKernelGGL((tmk::k_rows_r2c<N, NT>), dim3(nrows, nb), dim3(NT), 0, d.s, src, sbs, srs, 0, 0, N, 0, dst, dbs,
                     drs, twn, twh, 1.0f);
  chk("k_rows_r2c");
}
void rows_r2c(Dev& d, int N, const float* src, size_t sbs, int srs, cf* dst, size_t dbs, int drs, int nrows, int nb) {
  switch (N) {
#define TM_CASE(n) \
  case n: rows_r2c_t<n>(d, src, sbs, srs, dst, dbs, drs, nrows, nb); break;
    TM_LINE_SIZES(TM_CASE)
#undef TM_CASE
    default: throw PErr{TM_ERR_UNSUPPORTED, "unsupported row length " + std::to_string(N)};
  }
}
template <int N>
void cols_t(Dev& d, cf* data, int ncols, size_t bstride, int nb, int sign) {
  constexpr int NT = N >= 2048 ? 256 : 64;
  cf* tw = d.twiddles(N);
  if (sign < 0)
    hipLaunchKernelGGL((tmk::k_cols<N, NT, -1, 0>), dim3(ncols * nb), dim3(NT), 0, d.s, data, ncols, bstride, tw,
                       (const cf*)nullptr, 1.0f);
  else
    hipLaunchKernelGGL((tmk::k_cols<N, NT, +1, 0>), dim3(ncols * nb), dim3(NT), 0, d.s, data, ncols, bstride, tw,
                       (const cf*)nullptr, 1.0f);
  chk("k_cols");
}
void cols(Dev& d, int N, cf* data, int ncols, size_t bstride, int nb, int sign = -1) {
  switch (N) {
#define TM_CASE(n) \
  case n: cols_t<n>(d, data, ncols, bstride, nb, sign); break;
    TM_LINE_SIZES(TM_CASE)
#undef TM_CASE
    default: throw PErr{TM_ERR_UNSUPPORTED, "unsupported column length " + std::to_string(N)};
  }
}
template <int N>
void rows_c2r_t(Dev& d, const cf* in, size_t ibs, int irs, float* out, size_t obs, int ors, int nrows, int nb) {
  constexpr int NT = (N / 2) >= 1024 ? 256 : 64;
  hipLaunchKernelGGL((tmk::k_rows_c2r<N, NT>), dim3(nrows, nb), dim3(NT), 0, d.s, in, ibs, irs, out, obs, ors,
                     d.twiddles(N), d.twiddles(N / 2));
  chk("k_rows_c2r");
}
void rows_c2r(Dev& d, int N, const cf* in, size_t ibs, int irs, float* out, size_t obs, int ors, int nrows, int nb) {
  switch (N) {
#define TM_CASE(n) \
  case n: rows_c2r_t<n>(d, in, ibs, irs, out, obs, ors, nrows, nb); break;
    TM_LINE_SIZES(TM_CASE)
#undef TM_CASE
    default: throw PErr{TM_ERR_UNSUPPORTED, "unsupported row length " + std::to_string(N)};
  }
}

// Unscaled R2C of a real image (padded layout, nz planes) on the device; returns the spectrum.
std::vector<std::complex<float>> forward_fft(int device, const Img& a) {
  P_REQ(line_size_compiled(a.nx) && line_size_compiled(a.ny) && (a.nz == 1 || line_size_compiled(a.nz)),
        TM_ERR_UNSUPPORTED, "no compiled device FFT for size " + std::to_string(a.nx) + "x" + std::to_string(a.ny) +
                                "x" + std::to_string(a.nz));
  Dev d(device);
  const int nh = a.nx / 2 + 1;
  const size_t plane = (size_t)nh * a.ny;
  float* src = d.alloc<float>(a.v.size());
  cf* spec = d.alloc<cf>(plane * a.nz);
  P_HIP(hipMemcpy(src, a.v.data(), a.v.size() * sizeof(float), hipMemcpyHostToDevice));
  rows_r2c(d, a.nx, src, (size_t)a.rs() * a.ny, a.rs(), spec, plane, nh, a.ny, a.nz);
  cols(d, a.ny, spec, nh, plane, a.nz);
  if (a.nz > 1) cols(d, a.nz, spec, (int)plane, 0, 1);
  P_HIP(hipStreamSynchronize(d.s));
  std::vector<std::complex<float>> out(plane * a.nz);
  P_HIP(hipMemcpy(out.data(), spec, out.size() * sizeof(cf), hipMemcpyDeviceToHost));
  return out;
}

// Unscaled C2R of a 2-D spectrum (H x (W/2+1)) on the device (BackwardFFT, image.cpp:5104-5116);
// returns the real image in the padded layout.
Img backward_fft(int device, const std::vector<std::complex<float>>& spec, int nx, int ny) {
  P_REQ(line_size_compiled(nx) && line_size_compiled(ny), TM_ERR_UNSUPPORTED,
        "no compiled device FFT for size " + std::to_string(nx) + "x" + std::to_string(ny));
  Dev d(device);
  const int nh = nx / 2 + 1;
  cf* buf = d.alloc<cf>((size_t)nh * ny);
  P_HIP(hipMemcpy(buf, spec.data(), spec.size() * sizeof(cf), hipMemcpyHostToDevice));
  cols(d, ny, buf, nh, 0, 1, +1);
  rows_c2r(d, nx, buf, 0, nh, reinterpret_cast<float*>(buf), 0, nx + 2, ny, 1);
  P_HIP(hipStreamSynchronize(d.s));
  Img out;
  out.alloc(nx, ny, 1);
  P_HIP(hipMemcpy(out.v.data(), buf, out.v.size() * sizeof(float), hipMemcpyDeviceToHost));
  return out;
}

// ---------------------------------------------------------------- transforms of any length
// Bluestein tables for length N and sign s on the device: chirp c_n = exp(s i pi n^2 / N) (angles
// from n^2 mod 2N in long double) and bhat = FFT_M(b) / M, b_m = conj(c_|m|) for |m| < N.
struct Bluestein {
  int N = 0, M = 0;
  cf* chirp = nullptr;
  cf* bhat = nullptr;
};
Bluestein bluestein(Dev& d, int N, int sign) {
  Bluestein b;
  b.N = N;
  b.M = 1;
  while (b.M < 2 * N - 1) b.M *= 2;
  P_REQ(b.M == 2048 || b.M == 4096, TM_ERR_UNSUPPORTED, "transform length " + std::to_string(N) + " not supported");
  const long double pi = acosl(-1.0L);
  std::vector<cf> c(N), w(b.M, cf{0.f, 0.f});
  for (long n = 0; n < N; ++n) {
    const long double a = sign * pi * (long double)((n * n) % (2L * N)) / (long double)N;
    c[n].x = (float)cosl(a);
    c[n].y = (float)sinl(a);
  }
  for (int m = 0; m < N; ++m) {
    const cf cc{c[m].x, -c[m].y};
    w[m] = cc;
    if (m > 0) w[b.M - m] = cc;
  }
  b.chirp = d.alloc<cf>(N);
  b.bhat = d.alloc<cf>(b.M);
  P_HIP(hipMemcpy(b.chirp, c.data(), N * sizeof(cf), hipMemcpyHostToDevice));
  P_HIP(hipMemcpy(b.bhat, w.data(), b.M * sizeof(cf), hipMemcpyHostToDevice));
  cols(d, b.M, b.bhat, 1, 0, 1, -1);
  std::vector<cf> h(b.M);
  P_HIP(hipStreamSynchronize(d.s));
  P_HIP(hipMemcpy(h.data(), b.bhat, b.M * sizeof(cf), hipMemcpyDeviceToHost));
  const float inv = (float)(1.0 / b.M);
  for (auto& v : h) { v.x *= inv; v.y *= inv; }
  P_HIP(hipMemcpy(b.bhat, h.data(), b.M * sizeof(cf), hipMemcpyHostToDevice));
  return b;
}
// Complex transforms of length N and sign s down columns (k_cols layout), any N.
void cols_any(Dev& d, int N, cf* data, int ncols, size_t bstride, int nb, int sign) {
  if (line_size_compiled(N)) { cols(d, N, data, ncols, bstride, nb, sign); return; }
  Bluestein b = bluestein(d, N, sign);
  const dim3 g((unsigned)((size_t)ncols * nb));
  if (b.M == 2048)
    hipLaunchKernelGGL((tmk::k_bluestein_cols<2048, 256>), g, dim3(256), 0, d.s, data, N, ncols, bstride, b.chirp, b.bhat,
                       d.twiddles(2048));
  else
    hipLaunchKernelGGL((tmk::k_bluestein_cols<4096, 256>), g, dim3(256), 0, d.s, data, N, ncols, bstride, b.chirp, b.bhat,
                       d.twiddles(4096));
  chk("k_bluestein_cols");
}
// C2R rows of any even length N.
void c2r_rows_any(Dev& d, int N, const cf* in, size_t ibs, int irs, float* out, size_t obs, int ors, int nrows, int nb) {
  if (line_size_compiled(N)) { rows_c2r(d, N, in, ibs, irs, out, obs, ors, nrows, nb); return; }
  P_REQ(N % 2 == 0, TM_ERR_UNSUPPORTED, "odd C2R length");
  Bluestein b = bluestein(d, N, +1);
  if (b.M == 2048)
    hipLaunchKernelGGL((tmk::k_bluestein_c2r_rows<2048, 256>), dim3(nrows, nb), dim3(256), 0, d.s, in, N, ibs, irs, out, obs,
                       ors, b.chirp, b.bhat, d.twiddles(2048));
  else
    hipLaunchKernelGGL((tmk::k_bluestein_c2r_rows<4096, 256>), dim3(nrows, nb), dim3(256), 0, d.s, in, N, ibs, irs, out, obs,
                       ors, b.chirp, b.bhat, d.twiddles(4096));
  chk("k_bluestein_c2r_rows");
}

// Image::ReturnAverageOfRealValues with a mask radius (image.cpp:6263-6322): the mean of the voxels
// whose squared distance from the box centre is within the squared radius.
float masked_average(const Img& a, float radius) {
  double sum = 0.0;
  long n = 0;
  const float r2 = powf(radius, 2);
  for (int k = 0; k < a.nz; ++k) {
    const float z = powf(k - a.nz / 2, 2);
    for (int j = 0; j < a.ny; ++j) {
      const float y = powf(j - a.ny / 2, 2);
      for (int i = 0; i < a.nx; ++i) {
        const float x = powf(i - a.nx / 2, 2);
        if (x + y + z <= r2) { sum += a.at(i, j, k); ++n; }
      }
    }
  }
  return n > 0 ? float(sum / n) : 0.0f;
}

// Image::ChangePixelSize(other of the same size, factor, tolerance, return_fft = true)
// (image.cpp:8705-8792) of a real cube on the device: pad to the intermediate size with the masked
// average, forward transform (scaled), crop or pad the spectrum to the new size (with the Nyquist
// planes ClipInto fills in), inverse transform, clip back to the cube's size with the same padding
// value, forward transform (scaled). Returns the half spectrum.
std::vector<std::complex<float>> change_pixel_size_cube(int device, const Img& vol, float factor, float tol) {
  const int P = vol.nx;
  const int min_dimension = P;
  int pad, nw;
  float padding_value = 0.0f;
  if (tol < 1.0f / (min_dimension * factor)) {
    const float inter = 1.0f / (tol * min_dimension);
    pad = P * inter;
    if (even(P) != even(pad)) pad++;
    nw = int(pad / factor);
    if (even(P) != even(nw)) nw++;
  } else {
    pad = P * factor;
    if (even(P) != even(pad)) pad++;
    nw = P;
  }
  P_REQ(line_size_compiled(pad), TM_ERR_UNSUPPORTED,
        "pixel-size change pads to " + std::to_string(pad) + ", a length without a compiled transform");
  if (pad > P) padding_value = masked_average(vol, 0.45f * min_dimension);
  Dev d(device);
  // real cube, then the padded cube
  float* vd = d.alloc<float>(vol.v.size());
  P_HIP(hipMemcpy(vd, vol.v.data(), vol.v.size() * sizeof(float), hipMemcpyHostToDevice));
  const size_t pad_floats = (size_t)(pad + 2) * pad * pad;
  float* pb = d.alloc<float>(pad_floats);
  hipLaunchKernelGGL(tmk::k_real_clip3, dim3(8192), dim3(256), 0, d.s, vd, P, pb, pad, padding_value);
  chk("k_real_clip3");
  const int ph = pad / 2 + 1;
  cf* ps = d.alloc<cf>((size_t)ph * pad * pad);
  const float inv_pad = 1. / (float)((long)pad * pad * pad);
  rows_r2c(d, pad, pb, (size_t)(pad + 2) * pad, pad + 2, ps, (size_t)ph * pad, ph, pad, pad);
  P_HIP(hipStreamSynchronize(d.s));
  cols(d, pad, ps, ph, (size_t)ph * pad, pad);
  cols(d, pad, ps, ph * pad, 0, 1);
  // the ForwardFFT scale 1/N, applied to the spectrum
  hipLaunchKernelGGL(tmk::k_scale_cf, dim3(8192), dim3(256), 0, d.s, ps, (size_t)ph * pad * pad, inv_pad);
  chk("k_scale_cf");
  // Resize in Fourier space to the new size
  const int nh = nw / 2 + 1;
  const size_t new_floats = (size_t)(nw + 2) * nw * nw;
  float* nbf = d.alloc<float>(new_floats);
  cf* ns = reinterpret_cast<cf*>(nbf);
  hipLaunchKernelGGL(tmk::k_fourier_clip3, dim3(8192), dim3(256), 0, d.s, ps, pad, ns, nw);
  chk("k_fourier_clip3");
  if (nw > pad) {
    hipLaunchKernelGGL(tmk::k_fourier_grow_y, dim3((ph * pad + 255) / 256), dim3(256), 0, d.s, ps, pad, ns, nw);
    chk("k_fourier_grow_y");
    hipLaunchKernelGGL(tmk::k_fourier_grow_z, dim3((ph * (pad + 1) + 255) / 256), dim3(256), 0, d.s, ns, pad, nw);
    chk("k_fourier_grow_z");
  }
  std::vector<std::complex<float>> out;
  if (nw == P) {
    // the new size is the cube's: the cropped spectrum is the result (other_image->CopyFrom)
    P_HIP(hipStreamSynchronize(d.s));
    out.resize((size_t)nh * nw * nw);
    P_HIP(hipMemcpy(out.data(), ns, out.size() * sizeof(cf), hipMemcpyDeviceToHost));
    return out;
  }
  // BackwardFFT at the new size (z, y columns, then the C2R rows)
  cols_any(d, nw, ns, nh * nw, 0, 1, +1);
  cols_any(d, nw, ns, nh, (size_t)nh * nw, nw, +1);
  c2r_rows_any(d, nw, ns, (size_t)nh * nw, nh, nbf, (size_t)(nw + 2) * nw, nw + 2, nw, nw);
  // ClipInto the cube's size with the padding value, then ForwardFFT
  const int Ph = P / 2 + 1;
  float* cb = d.alloc<float>((size_t)(P + 2) * P * P);
  hipLaunchKernelGGL(tmk::k_real_clip3, dim3(8192), dim3(256), 0, d.s, nbf, nw, cb, P, padding_value);
  chk("k_real_clip3");
  cf* cs = d.alloc<cf>((size_t)Ph * P * P);
  rows_r2c(d, P, cb, (size_t)(P + 2) * P, P + 2, cs, (size_t)Ph * P, Ph, P, P);
  P_HIP(hipStreamSynchronize(d.s));
  cols(d, P, cs, Ph, (size_t)Ph * P, P);
  cols(d, P, cs, Ph * P, 0, 1);
  P_HIP(hipStreamSynchronize(d.s));
  out.resize((size_t)Ph * P * P);
  P_HIP(hipMemcpy(out.data(), cs, out.size() * sizeof(cf), hipMemcpyDeviceToHost));
  const float inv_n = 1. / (float)((long)P * P * P);
  for (auto& c : out) c *= inv_n;
  return out;
}

// Image::SwapRealSpaceQuadrants on a non-centred even-size spectrum (phase shift n/2 per axis,
// image.cpp:9221-9268, functions.h:494-504), applied in place.
void swap_quadrants(std::complex<float>* c, int nx, int ny, int nz) {
  const int nh = nx / 2 + 1;
  const float sx = (float)(nx / 2), sy = (float)(ny / 2), sz = nz > 1 ? (float)(nz / 2) : 0.0f;
  std::vector<float> px(nh), py(ny), pz(nz);
  for (int i = 0; i < nh; ++i) px[i] = sx * (float)i * 2.0 * PI / (float)nx;
  for (int j = 0; j < ny; ++j) py[j] = sy * (float)(j >= ny / 2 ? j - ny : j) * 2.0 * PI / (float)ny;
  for (int k = 0; k < nz; ++k) pz[k] = sz * (float)(k >= nz / 2 && nz > 1 ? k - nz : k) * 2.0 * PI / (float)nz;
  size_t a = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nh; ++i, ++a) {
        const float ph = -px[i] - py[j] - pz[k];
        c[a] *= std::complex<float>(cosf(ph), sinf(ph));
      }
}

// Curve (curve.cpp:247-259, 377-396, 504-560) restricted to what the whitening filter needs.
struct Curve {
  std::vector<float> x, y;
  void setup(float lo, float hi, int n) {
    x.resize(n); y.assign(n, 0.f);
    for (int c = 0; c < n; ++c) x[c] = c * (hi - lo) / float(n - 1) + lo;
  }
  void add(float wx, float val) {
    const int n = (int)x.size();
    const int idx = (x[0] == x[1]) ? n - 2 : int((wx - x[0]) / (x[1] - x[0]));
    if (idx == n - 1) { y[idx] += val; return; }
    const float d = (wx - x[idx]) / (x[idx + 1] - x[idx]);
    const float vm = val * (1.0 - d), vn = val * d;
    y[idx] += vm;
    y[idx + 1] += vn;
  }
  float interp(float wx) const {
    const int n = (int)x.size();
    int i = n - 1;
    if (wx < x[0]) i = 0;
    else {
      // unique bin with x[c] <= wx < x[c+1]; equal spacing gives it directly, checked on both sides
      int c = int((wx - x[0]) / (x[1] - x[0]));
      c = std::max(0, std::min(c, n - 1));
      while (c > 0 && wx < x[c]) --c;
      while (c < n - 1 && wx >= x[c + 1]) ++c;
      i = c;
    }
    if (i == n - 1) return y[n - 1];
    float r = 0.0;
    const float d = (wx - x[i]) / (x[i + 1] - x[i]);
    r += y[i] * (1.0 - d);
    r += y[i + 1] * d;
    return r;
  }
};

// Image::ApplyCurveFilter (image.cpp:7284-7313) on a 2-D spectrum (resolution limit 1)
void apply_curve(std::complex<float>* c, int nx, int ny, const Curve& f) {
  const int nh = nx / 2 + 1;
  for (int j = 0; j < ny; ++j) {
    const float sy = std::pow((j >= (even(ny) ? ny / 2 : ny / 2 + 1) ? j - ny : j) * (1.0 / double(ny)), 2);
    for (int i = 0; i < nh; ++i) {
      const float sx = std::pow(i * (1.0 / double(nx)), 2);
      const float s2 = sx + sy + 0.0f;
      const float sf = std::sqrt((double)s2);
      if (sf <= 1.0f) c[(size_t)j * nh + i] *= f.interp(sf);
      else c[(size_t)j * nh + i] = 0.f;
    }
  }
}

// Image::ReturnSumOfSquares, Fourier branch (image.cpp:290-311), 2-D
float spectrum_sum_of_squares(const std::complex<float>* c, int nx, int ny) {
  const int nh = nx / 2 + 1;
  const int fny = even(ny) ? ny / 2 : ny / 2 + 1;
  const int lo_y = even(ny) ? -ny / 2 : -(ny - 1) / 2;
  const bool xe = even(nx);
  double sum = 0.0;
  size_t a = 0;
  for (int j = 0; j < ny; ++j) {
    const int jj = j >= fny ? j - ny : j;
    for (int i = 0; i < nh; ++i, ++a) {
      const double a2 = std::pow((double)std::abs(c[a]), 2);
      const bool edge_x = (i == 0 || (i == nx / 2 && xe));
      if (edge_x && (jj == 0 || (jj == lo_y && xe))) sum += a2 * 0.5;
      else if ((i != 0 && (i != nx / 2 || !xe)) || jj >= 0) sum += a2;
    }
  }
  return 2.0 * sum;
}

// ReplaceOutliersWithMean(5) (image.cpp:986-1100) over the logical pixels of a 2-D image
void replace_outliers(Img& img) {
  const int W = img.nx, H = img.ny;
  double s = 0.0, s2 = 0.0;
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) { const float v = img.at(x, y); s += v; s2 += powf(v, 2); }
  const long n = (long)W * H;
  const float sigma = sqrtf(fabsf(float(s2 / n - powf(s / n, 2))));
  const float mean = float(s / n);
  const float mx_ = mean + 5.0f * sigma, mn_ = mean - 5.0f * sigma;
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      float& v = img.at(x, y);
      if (v > mx_) v = mean;
      else if (v < mn_) v = mean;
    }
}

// Whitening curve of a spectrum (H x (W/2+1), DC already 0): Compute1DPowerSpectrumCurve
// (image.cpp:7135-7205; intensities, Hermitian mates skipped), SquareRoot, Reciprocal, scaled to a
// maximum of 1 (match_template.cpp:711-716, refine_template.cpp:536-540).
Curve whitening_curve(const std::complex<float>* spec, int W, int H, int ncurve) {
  Curve wf, nt;
  wf.setup(0.0, 0.5 * sqrtf(2.0), ncurve);
  nt.setup(0.0, 0.5 * sqrtf(2.0), ncurve);
  const int nh = W / 2 + 1;
  const int fny = H / 2;
  for (int j = 0; j < H; ++j) {
    const float sy = powf((j >= fny ? j - H : j) * (1.0 / double(H)), 2);
    for (int i = 0; i < nh; ++i) {
      if (i == 0 && j >= fny) continue;
      const float sx = powf(i * (1.0 / double(W)), 2);
      const float sf = sqrtf(sx + sy + 0.0f);
      const std::complex<float> c = spec[(size_t)j * nh + i];
      wf.add(sf, c.real() * c.real() + c.imag() * c.imag());
      nt.add(sf, 1.0);
    }
  }
  for (int c = 0; c < ncurve; ++c) wf.y[c] = nt.y[c] > 0.0 ? wf.y[c] / nt.y[c] : 0.0f;
  for (auto& a : wf.y) a = sqrtf(a);
  for (auto& a : wf.y) if (a != 0.0) a = 1.0 / a;
  float mxv = -FLT_MAX;
  for (float a : wf.y) mxv = std::max(mxv, a);
  const float k = 1.0f / mxv;
  for (auto& a : wf.y) a *= k;
  return wf;
}

}  // namespace

extern "C" {

const char* tm_prepare_last_error(void) { return g_perr.c_str(); }

int tm_prepare_sizes(int mx, int my, int template_size, int* W, int* H, int* rotated, int* n_curve, int* max_padding) {
  return pguard([&] {
    P_REQ(mx > 0 && my > 0 && template_size > 0 && W && H && rotated && n_curve, TM_ERR_INVALID, "bad arguments");
    int fx = factorize_axis(mx, template_size);
    int fy = factorize_axis(my, template_size);
    int mp = std::max(0, std::max(fx - mx, fy - my));
    int rot = (!is_power_of_two(fx) && is_power_of_two(fy)) ? 1 : 0;
    *W = rot ? fy : fx;
    *H = rot ? fx : fy;
    *rotated = rot;
    *n_curve = int((*W / 2.0 + 1.0) * sqrtf(2.0) + 1.0);
    if (max_padding) *max_padding = mp;
  });
}

int tm_prepare_input_image(int device, const float* micrograph, int mx, int my, int template_size, float* spectrum,
                           float* curve_x, float* curve_y) {
  return pguard([&] {
    P_REQ(micrograph && spectrum && curve_x && curve_y, TM_ERR_INVALID, "NULL argument");
    int W, H, rot, ncurve, mp;
    if (tm_prepare_sizes(mx, my, template_size, &W, &H, &rot, &ncurve, &mp) != TM_OK) throw PErr{TM_ERR_INVALID, g_perr};
    Img img;
    img.alloc(mx, my, 1);
    for (int y = 0; y < my; ++y) std::memcpy(&img.at(0, y), micrograph + (size_t)y * mx, mx * sizeof(float));
    const int fx = rot ? H : W, fy = rot ? W : H;
    if (fx != mx || fy != my) img = clip_into(img, fx, fy, 1, edge_mean(img));
    if (rot) {  // Image::Rotate2DInPlaceBy90Degrees(true) (image.cpp:11104-11141)
      Img b;
      b.alloc(img.ny, img.nx, 1);
      for (int yo = 0; yo < img.ny; ++yo)
        for (int xo = 0; xo < img.nx; ++xo) b.at(img.ny - yo - 1, xo) = img.at(xo, yo);
      img = std::move(b);
    }
    replace_outliers(img);
    auto spec = forward_fft(device, img);
    const float inv_n = 1. / (float)((long)W * H);  // ForwardFFT scaling (image.cpp:5096-5099)
    for (auto& c : spec) c *= inv_n;
    swap_quadrants(spec.data(), W, H, 1);
    spec[0] = 0.f;
    Curve wf = whitening_curve(spec.data(), W, H, ncurve);
    apply_curve(spec.data(), W, H, wf);
    spec[0] = 0.f;
    const float inv = 1. / sqrtf(spectrum_sum_of_squares(spec.data(), W, H));
    for (auto& c : spec) c *= inv;
    std::memcpy(spectrum, spec.data(), spec.size() * sizeof(cf));
    std::memcpy(curve_x, wf.x.data(), ncurve * sizeof(float));
    std::memcpy(curve_y, wf.y.data(), ncurve * sizeof(float));
  });
}

int tm_prepare_template(int device, const float* volume, int t, float padding, float* spectrum, int* P_out) {
  return pguard([&] {
    P_REQ(volume && t > 0, TM_ERR_INVALID, "bad arguments");
    const int P = padding != 1.0f ? int(t * padding) : t;
    if (P_out) *P_out = P;
    if (!spectrum) return;
    Img v;
    v.alloc(t, t, t);
    for (int z = 0; z < t; ++z)
      for (int y = 0; y < t; ++y) std::memcpy(&v.at(0, y, z), volume + ((size_t)z * t + y) * t, t * sizeof(float));
    if (P != t) v = clip_into(v, P, P, P, edge_mean(v));
    auto spec = forward_fft(device, v);
    const float inv_n = 1. / (float)((long)P * P * P);
    for (auto& c : spec) c *= inv_n;
    spec[0] = 0.f;
    swap_quadrants(spec.data(), P, P, P);
    std::memcpy(spectrum, spec.data(), spec.size() * sizeof(cf));
  });
}

int tm_prepare_template_scaled(int device, const float* volume, int t, float padding, float pixel_size_factor,
                               float* spectrum, int* P_out) {
  return pguard([&] {
    P_REQ(volume && t > 0 && pixel_size_factor > 0.0f, TM_ERR_INVALID, "bad arguments");
    const int P = padding != 1.0f ? int(t * padding) : t;
    if (P_out) *P_out = P;
    if (!spectrum) return;
    Img v;
    v.alloc(t, t, t);
    for (int z = 0; z < t; ++z)
      for (int y = 0; y < t; ++y) std::memcpy(&v.at(0, y, z), volume + ((size_t)z * t + y) * t, t * sizeof(float));
    if (P != t) v = clip_into(v, P, P, P, edge_mean(v));
    std::vector<std::complex<float>> spec;
    if (fabsf(pixel_size_factor - 1.0f) < 0.001f) {
      spec = forward_fft(device, v);
      const float inv_n = 1. / (float)((long)P * P * P);
      for (auto& c : spec) c *= inv_n;
    } else {
      spec = change_pixel_size_cube(device, v, pixel_size_factor, 0.001f);
    }
    spec[0] = 0.f;
    swap_quadrants(spec.data(), P, P, P);
    std::memcpy(spectrum, spec.data(), spec.size() * sizeof(cf));
  });
}

int tm_projection_filter(int T, const float* curve_x, const float* curve_y, int n_curve, float voltage_kV,
                         float spherical_aberration_mm, float amplitude_contrast, float defocus1, float defocus2,
                         float defocus_angle_deg, float pixel_size, float phase_shift_deg, float* out) {
  return pguard([&] {
    P_REQ(T > 0 && curve_x && curve_y && n_curve > 1 && out, TM_ERR_INVALID, "bad arguments");
    // CTF::Init (ctf.cpp:93-150) and SetDefocus (ctf.cpp:378-383)
    const float wl = 12.2639f / sqrtf(1000.0f * voltage_kV + 0.97845e-6 * powf(1000.0f * voltage_kV, 2)) / pixel_size;
    const float wl2 = powf(wl, 2);
    const float cs = spherical_aberration_mm * 10000000.0 / pixel_size;
    const float act = (fabs(amplitude_contrast - 1.0) < 1e-3)
                          ? (float)(PI / 2.0f)
                          : atanf(amplitude_contrast / sqrtf(1.0 - powf(amplitude_contrast, 2)));
    const float d1 = defocus1 / pixel_size, d2 = defocus2 / pixel_size;
    const float astig = deg_2_rad(defocus_angle_deg);
    const float phase = deg_2_rad(phase_shift_deg);
    Curve f;
    f.x.assign(curve_x, curve_x + n_curve);
    f.y.assign(curve_y, curve_y + n_curve);
    const int nh = T / 2 + 1;
    std::complex<float>* o = reinterpret_cast<std::complex<float>*>(out);
    for (int j = 0; j < T; ++j) {
      const float y2 = (j >= T / 2 ? j - T : j) * (1.0 / double(T));
      const float ysq = powf(y2, 2);
      for (int i = 0; i < nh; ++i) {
        const float x2 = i * (1.0 / double(T));
        const float az = (i == 0 && j == 0) ? 0.0f : atan2f(y2, x2);
        const float g2 = powf(x2, 2) + ysq;
        float ctf;
        if (d1 == 0.0f && d2 == 0.0f) ctf = -0.7f;
        else {
          const float df = 0.5f * (d1 + d2 + cosf(2.0f * (az - astig)) * (d1 - d2));
          ctf = -sinf(PIf * wl * g2 * (df - 0.5f * wl2 * g2 * cs) + phase + act);
        }
        o[(size_t)j * nh + i] = std::complex<float>(ctf, 0.f);
      }
    }
    apply_curve(o, T, T, f);
  });
}

int tm_search_grid(char symmetry_type, int symmetry_number, float angular_step, float* out, int capacity) {
  float phi_max, theta_max;
  bool test_mirror;
  switch (symmetry_type) {
    case 'C': phi_max = 360.0 / symmetry_number; theta_max = 90.0; test_mirror = true; break;
    case 'D': phi_max = 360.0 / symmetry_number; theta_max = 90.0; test_mirror = false; break;
    case 'T': phi_max = 180.0; theta_max = 54.7; test_mirror = false; break;
    case 'O': phi_max = 90.0; theta_max = 54.7; test_mirror = false; break;
    case 'I': phi_max = 180.0; theta_max = 31.7; test_mirror = false; break;
    default: g_perr = "invalid symmetry"; return TM_ERR_INVALID;
  }
  if (symmetry_type == 'C' && symmetry_number == 1 && test_mirror) theta_max = 180.0f;  // match_template.cpp:689-695
  const float theta_step = theta_max / int(theta_max / angular_step + 0.5);
  int n = 0;
  for (float theta = 0.0f; theta < theta_max + theta_step / 2.0; theta += theta_step) {
    float phi_step;
    if (theta == 0.0 || theta == 180.0) phi_step = phi_max;
    else {
      phi_step = fabsf(angular_step / sinf(deg_2_rad(theta)));
      if (phi_step > phi_max) phi_step = phi_max;
      phi_step = phi_max / int(phi_max / phi_step + 0.5);
    }
    for (float phi = 0.0; phi < phi_max; phi += phi_step) {
      if (out && n < capacity) { out[2 * n] = phi; out[2 * n + 1] = theta; }
      ++n;
    }
  }
  return n;
}

int tm_psi_list(float psi_start, float psi_max, float psi_step, float* out, int capacity) {
  int n = 0;
  for (float p = psi_start; p <= psi_max; p += psi_step) {
    if (out && n < capacity) out[n] = p;
    ++n;
  }
  return n;
}


int tm_prepare_refine_image(int device, const float* micrograph, int mx, int my, float* whitened, float* curve_x,
                            float* curve_y, int* n_curve) {
  return pguard([&] {
    P_REQ(micrograph && mx > 0 && my > 0 && even(mx), TM_ERR_INVALID, "bad arguments (the image width must be even)");
    const int ncurve = int((mx / 2.0 + 1.0) * sqrtf(2.0) + 1.0);
    if (n_curve) *n_curve = ncurve;
    if (!whitened) return;
    P_REQ(curve_x && curve_y, TM_ERR_INVALID, "NULL curve argument");
    Img img;
    img.alloc(mx, my, 1);
    for (int y = 0; y < my; ++y) std::memcpy(&img.at(0, y), micrograph + (size_t)y * mx, mx * sizeof(float));
    replace_outliers(img);
    auto spec = forward_fft(device, img);
    const float inv_n = 1. / (float)((long)mx * my);  // ForwardFFT scaling (image.cpp:5096-5099)
    for (auto& c : spec) c *= inv_n;
    spec[0] = 0.f;
    Curve wf = whitening_curve(spec.data(), mx, my, ncurve);
    apply_curve(spec.data(), mx, my, wf);
    spec[0] = 0.f;
    // DivideByConstant(sqrt(ReturnSumOfSquares())) (refine_template.cpp:544)
    const float inv = 1. / (float)sqrt((double)spectrum_sum_of_squares(spec.data(), mx, my));
    for (auto& c : spec) c *= inv;
    Img out = backward_fft(device, spec, mx, my);
    for (int y = 0; y < my; ++y) std::memcpy(whitened + (size_t)y * mx, &out.at(0, y), mx * sizeof(float));
    std::memcpy(curve_x, wf.x.data(), ncurve * sizeof(float));
    std::memcpy(curve_y, wf.y.data(), ncurve * sizeof(float));
  });
}

}  // extern "C"
