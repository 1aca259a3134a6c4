// TEST INFRASTRUCTURE ONLY — part of the CPU oracle for the template-matching path.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this code,
// and only as the checker; the product never links or calls it.
//
// A restatement of the subset of cisTEM's `Image` class (src/core/image.{h,cpp}) that the
// match_template hot path touches, templated on the real type R so that the same code
// gives the FP32 oracle (R = float, the reference's arithmetic) and an FP64 variant
// (R = double) used to measure precision. Every method cites the reference lines it follows;
// mixed float/double expressions of the reference are kept where they affect rounding.
#pragma once

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <complex>
#include <stdexcept>
#include <thread>
#include <vector>

#include "tm_oracle_fft.h"

namespace tmo {

// Worker threads of the oracle's large transforms (1 = serial). The result does not depend on it:
// every line is transformed independently.
inline int& fft_threads() {
  static int n = 1;
  return n;
}
template <class F>
void parallel_for(long n, F f) {
  const int nt = (int)std::min<long>(std::max(1, fft_threads()), std::max(1L, n));
  if (nt <= 1) { for (long i = 0; i < n; ++i) f(i); return; }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&, t] { for (long i = n * t / nt; i < n * (t + 1) / nt; ++i) f(i); });
  for (auto& x : th) x.join();
}

static const double PI = 3.14159265359;   // src/core/defines.h:6
static const float PIf = 3.14159265359f;  // src/core/defines.h:7

// src/core/functions.h:428-436 (float in, double arithmetic, float out)
inline float rad_2_deg(float radians) { return radians / (PIf / 180.); }
inline float deg_2_rad(float degrees) { return degrees * PIf / 180.; }
// src/core/functions.h:470-483
inline int myroundint(double a) { if (a > 0) return int(a + 0.5); else return int(a - 0.5); }
inline int myroundint(float a) { if (a > 0) return int(a + 0.5); else return int(a - 0.5); }
inline bool IsEven(int n) { return (n & 1) == 0; }
// src/core/functions.h:422-426
inline bool is_power_of_two(int n) {
  if (n == 0) return false;
  return (std::ceil(std::log2((float)n)) == std::floor(std::log2((float)n)));
}
// src/core/functions.h:494-504
inline float ReturnPhaseFromShift(float real_space_shift, float distance_from_origin, float dimension_size) {
  return real_space_shift * distance_from_origin * 2.0 * PI / dimension_size;
}
inline std::complex<float> Return3DPhaseFromIndividualDimensions(float phase_x, float phase_y, float phase_z) {
  float temp_phase = -phase_x - phase_y - phase_z;
  return std::complex<float>(cosf(temp_phase), sinf(temp_phase));
}

template <class R> struct Curve;
template <class R> struct CTFModel;

template <class R>
struct Image {
  typedef std::complex<R> C;
  int nx = 0, ny = 0, nz = 0;
  long real_memory_allocated = 0;
  int padding_jump_value = 2;
  long number_of_real_space_pixels = 0;
  bool is_in_real_space = true;
  bool object_is_centred_in_box = true;
  std::vector<R> v;

  // --- layout (image.cpp:4761-4846, 4933-5035) ---
  void Allocate(int x, int y, int z, bool real_space = true) {
    nx = x; ny = y; nz = z;
    long first = IsEven(x) ? x / 2 + 1 : (x - 1) / 2 + 1;
    real_memory_allocated = first * (long)y * (long)z * 2;
    v.assign(real_memory_allocated, (R)0);
    padding_jump_value = IsEven(nx) ? 2 : 1;
    number_of_real_space_pixels = long(nx) * long(ny) * long(nz);
    is_in_real_space = real_space;
    object_is_centred_in_box = true;
  }
  C* c() { return reinterpret_cast<C*>(v.data()); }
  const C* c() const { return reinterpret_cast<const C*>(v.data()); }
  int pubx() const { return nx / 2; }
  int puby() const { return ny - 1; }
  int pubz() const { return nz - 1; }
  int box_cx() const { return nx / 2; }
  int box_cy() const { return ny / 2; }
  int box_cz() const { return nz / 2; }
  int first_neg_y() const { return IsEven(ny) ? ny / 2 : ny / 2 + 1; }
  int first_neg_z() const { return IsEven(nz) ? nz / 2 : nz / 2 + 1; }
  int lcx_lo() const { return IsEven(nx) ? -nx / 2 : -(nx - 1) / 2; }
  int lcx_hi() const { return IsEven(nx) ? nx / 2 : (nx - 1) / 2; }
  int lcy_lo() const { return IsEven(ny) ? -ny / 2 : -(ny - 1) / 2; }
  int lcy_hi() const { return IsEven(ny) ? ny / 2 - 1 : (ny - 1) / 2; }
  int lcz_lo() const { return IsEven(nz) ? -nz / 2 : -(nz - 1) / 2; }
  int lcz_hi() const { return IsEven(nz) ? nz / 2 - 1 : (nz - 1) / 2; }
  double fvx() const { return 1.0 / double(nx); }
  double fvy() const { return 1.0 / double(ny); }
  double fvz() const { return 1.0 / double(nz); }
  long real_addr(int x, int y, int z) const {
    return (long(nx + padding_jump_value) * long(ny)) * long(z) + long(nx + padding_jump_value) * long(y) + long(x);
  }
  long four_addr_phys(int x, int y, int z) const {
    return (long(pubx() + 1) * long(puby() + 1)) * long(z) + long(pubx() + 1) * long(y) + long(x);
  }
  // image.h:218-272
  long four_addr_logical(int wx, int wy, int wz) const {
    int px, py, pz;
    if (wx >= 0) {
      px = wx;
      py = wy >= 0 ? wy : ny + wy;
      pz = wz >= 0 ? wz : nz + wz;
    } else {
      px = -wx;
      py = wy > 0 ? ny - wy : -wy;
      pz = wz > 0 ? nz - wz : -wz;
    }
    return four_addr_phys(px, py, pz);
  }
  // image.cpp:6899-6921
  int logical_y(int j) const { return j >= first_neg_y() ? j - ny : j; }
  int logical_z(int k) const { return k >= first_neg_z() ? k - nz : k; }

  // --- FFT (image.cpp:5085-5116) ---
  void ForwardFFT(bool should_scale = true) {
    fft_r2c_inplace();
    is_in_real_space = false;
    if (should_scale) DivideByConstant(float(number_of_real_space_pixels));
  }
  void BackwardFFT() {
    fft_c2r_inplace();
    is_in_real_space = true;
  }
  void fft_r2c_inplace() {
    if (!IsEven(nx)) throw std::runtime_error("oracle FFT needs even x dimension");
    const int hc = nx / 2 + 1;
    parallel_for((long)nz * ny, [&](long r) {
      std::vector<C> scratch(nx + 2);
      rfft_row(&v[real_addr(0, (int)(r % ny), (int)(r / ny))], nx, scratch.data());
    });
    if (ny > 1) lines_c2c(ny, (long)hc, /*dist*/ 1, /*count per plane*/ hc, (long)hc * ny, -1);
    if (nz > 1) lines_c2c(nz, (long)hc * ny, 1, hc * ny, 0, -1);
  }
  void fft_c2r_inplace() {
    if (!IsEven(nx)) throw std::runtime_error("oracle FFT needs even x dimension");
    const int hc = nx / 2 + 1;
    if (nz > 1) lines_c2c(nz, (long)hc * ny, 1, hc * ny, 0, +1);
    if (ny > 1) lines_c2c(ny, (long)hc, 1, hc, (long)hc * ny, +1);
    parallel_for((long)nz * ny, [&](long r) {
      std::vector<C> scratch(nx + 2);
      irfft_row(&v[real_addr(0, (int)(r % ny), (int)(r / ny))], nx, scratch.data());
    });
  }
  // Complex FFTs of length n along a strided axis. Lines: for each plane p (only when
  // plane_stride != 0, nz planes) and each of `count` lines starting at offset `l*dist`,
  // elements at stride `stride`. Lines are gathered in blocks of 8 for cache locality.
  void lines_c2c(int n, long stride, long dist, int count, long plane_stride, int sign) {
    auto plan = fft_plan<R>(n);
    const int planes = plane_stride ? nz : 1;
    const int BL = 8;
    const long blocks = (count + BL - 1) / BL;
    C* d = c();
    parallel_for((long)planes * blocks, [&](long q) {
      std::vector<C> buf((size_t)BL * n), scratch(n);
      const long p = q / blocks;
      const int l0 = (int)(q % blocks) * BL;
      C* base = d + p * plane_stride;
      const int nb = std::min(BL, count - l0);
      for (int e = 0; e < n; ++e) {
        C* src = base + (long)e * stride + (long)l0 * dist;
        for (int b = 0; b < nb; ++b) buf[(size_t)b * n + e] = src[b * dist];
      }
      for (int b = 0; b < nb; ++b) fft_c2c(*plan, &buf[(size_t)b * n], scratch.data(), sign);
      for (int e = 0; e < n; ++e) {
        C* dst = base + (long)e * stride + (long)l0 * dist;
        for (int b = 0; b < nb; ++b) dst[b * dist] = buf[(size_t)b * n + e];
      }
    });
  }

  // --- simple arithmetic (image.cpp:5118-5127, 5633-5641, 5188-5196, 6498-6525) ---
  void DivideByConstant(float constant_to_divide_by) {
    float inverse = 1. / constant_to_divide_by;
    for (long i = 0; i < real_memory_allocated; ++i) v[i] *= (R)inverse;
  }
  void MultiplyByConstant(float k) {
    for (long i = 0; i < real_memory_allocated; ++i) v[i] *= (R)k;
  }
  void AddConstant(float k) {
    for (long i = 0; i < real_memory_allocated; ++i) v[i] += (R)k;
  }
  void SetToConstant(float k) { std::fill(v.begin(), v.end(), (R)k); }
  void ZeroCentralPixel() {
    if (!is_in_real_space) { c()[0] = C(0, 0); return; }
    long address = 0;
    for (int j = 0; j < ny; ++j) {
      for (int i = 0; i < nx; ++i) {
        if (j == box_cy() && i == box_cx()) v[address] = 0;
        address++;
      }
      address += padding_jump_value;
    }
  }
  // image.cpp:829-891 (Fourier x Fourier branch)
  void MultiplyPixelWise(const Image& o) {
    if (is_in_real_space) {
      for (long i = 0; i < real_memory_allocated; ++i) v[i] *= o.v[i];
    } else {
      C* a = c();
      const C* b = o.c();
      for (long i = 0; i < real_memory_allocated / 2; ++i) a[i] *= b[i];
    }
  }

  // --- phase shift / quadrant swap (image.cpp:9221-9268, 9757-9822) ---
  void PhaseShift(float sx, float sy, float sz) {
    bool need_to_fft = false;
    if (is_in_real_space) { ForwardFFT(); need_to_fft = true; }
    long pc = 0;
    C* d = c();
    for (int k = 0; k <= pubz(); ++k) {
      float phase_z = ReturnPhaseFromShift(sz, (float)logical_z(k), (float)nz);
      for (int j = 0; j <= puby(); ++j) {
        float phase_y = ReturnPhaseFromShift(sy, (float)logical_y(j), (float)ny);
        for (int i = 0; i <= pubx(); ++i) {
          float phase_x = ReturnPhaseFromShift(sx, (float)i, (float)nx);
          std::complex<float> ph = Return3DPhaseFromIndividualDimensions(phase_x, phase_y, phase_z);
          d[pc] *= C((R)ph.real(), (R)ph.imag());
          pc++;
        }
      }
    }
    if (need_to_fft) BackwardFFT();
  }
  void SwapRealSpaceQuadrants() {
    bool must_fft = false;
    if (is_in_real_space) { must_fft = true; ForwardFFT(); }
    float xs, ys, zs;
    if (object_is_centred_in_box) {
      xs = float(box_cx()); ys = float(box_cy()); zs = float(box_cz());
    } else {
      xs = IsEven(nx) ? float(box_cx()) : float(box_cx()) - 1.0;
      ys = IsEven(ny) ? float(box_cy()) : float(box_cy()) - 1.0;
      zs = IsEven(nz) ? float(box_cz()) : float(box_cz()) - 1.0;
    }
    if (nz == 1) zs = 0.0;
    PhaseShift(xs, ys, zs);
    if (must_fft) BackwardFFT();
    object_is_centred_in_box = !object_is_centred_in_box;
  }

  // --- reductions ---
  // image.cpp:185-311 (no mask)
  float ReturnSumOfSquares() const {
    double sum = 0.0;
    long address = 0;
    if (is_in_real_space) {
      for (int k = 0; k < nz; ++k) {
        for (int j = 0; j < ny; ++j) {
          for (int i = 0; i < nx; ++i) { sum += std::pow((double)v[address], 2); address++; }
          address += padding_jump_value;
        }
      }
      return sum / nx / ny / nz;
    }
    const bool x_is_even = IsEven(nx);
    const C* d = c();
    for (int k = 0; k <= pubz(); ++k) {
      int kk = logical_z(k);
      for (int j = 0; j <= puby(); ++j) {
        int jj = logical_y(j);
        for (int i = 0; i <= pubx(); ++i) {
          double a2 = std::pow((double)std::abs(d[address]), 2);
          if ((i == 0 || (i == lcx_hi() && x_is_even)) && (jj == 0 || (jj == lcy_lo() && x_is_even)) &&
              (kk == 0 || (kk == lcz_lo() && x_is_even)))
            sum += a2 * 0.5;
          else if ((i == 0 || (i == lcx_hi() && x_is_even)) && nz != 1)
            sum += a2 * 0.25;
          else if ((i != 0 && (i != lcx_hi() || !x_is_even)) || (jj >= 0 && kk >= 0))
            sum += a2;
          address++;
        }
      }
    }
    return 2.0 * sum;
  }
  // image.cpp:6263-6340 with a mask radius: the pixels whose squared distance from the box centre is
  // within the squared radius
  float ReturnAverageOfRealValues(float wanted_mask_radius) const {
    if (!(wanted_mask_radius > 0.0)) return ReturnAverageOfRealValues();
    double sum = 0.0;
    long address = 0, number_of_pixels = 0;
    const float mask_radius_squared = powf(wanted_mask_radius, 2);
    for (int k = 0; k < nz; ++k) {
      const float z = powf(k - box_cz(), 2);
      for (int j = 0; j < ny; ++j) {
        const float y = powf(j - box_cy(), 2);
        for (int i = 0; i < nx; ++i) {
          const float x = powf(i - box_cx(), 2);
          const float d2 = x + y + z;
          if (d2 <= mask_radius_squared) { sum += v[address]; number_of_pixels++; }
          address++;
        }
        address += padding_jump_value;
      }
    }
    return number_of_pixels > 0 ? float(sum / number_of_pixels) : 0.0f;
  }
  // Image::CosineMask (image.cpp:3934-4209), real space, object centred in the box, not inverted, mask
  // value not forced (the call of refine_template.cpp:716): the pixels whose squared distance from
  // the box centre lies in [r0^2, (r0 + edge)^2], r0 = radius - edge / 2 (at least 0), are averaged
  // (double sum); inside r0 the pixels are kept, in the band they become v e + (1 - e) mean with
  // e = (1 + cos(pi (d - r0) / edge)) / 2, outside they become the mean. Returns the mask volume.
  float CosineMask(float wanted_mask_radius, float wanted_mask_edge) {
    if (!is_in_real_space || !object_is_centred_in_box)
      throw std::runtime_error("oracle CosineMask: only the centred real-space branch is restated");
    float mask_radius = wanted_mask_radius - wanted_mask_edge * 0.5;
    if (mask_radius < 0.0) mask_radius = 0.0;
    const float mask_radius_plus_edge = mask_radius + wanted_mask_edge;
    const float r2 = powf(mask_radius, 2), re2 = powf(mask_radius_plus_edge, 2);
    double pixel_sum = 0.0, mask_volume = 0.0;
    long number_of_pixels = 0, address = 0;
    for (int k = 0; k < nz; ++k) {
      const float z = powf(k - box_cz(), 2);
      for (int j = 0; j < ny; ++j) {
        const float y = powf(j - box_cy(), 2);
        for (int i = 0; i < nx; ++i) {
          const float d2 = powf(i - box_cx(), 2) + y + z;
          if (d2 >= r2 && d2 <= re2) { pixel_sum += v[address]; number_of_pixels++; }
          address++;
        }
        address += padding_jump_value;
      }
    }
    if (number_of_pixels == 0) throw std::runtime_error("oracle CosineMask: no pixels in the edge band");
    pixel_sum /= number_of_pixels;
    address = 0;
    for (int k = 0; k < nz; ++k) {
      const float z = powf(k - box_cz(), 2);
      for (int j = 0; j < ny; ++j) {
        const float y = powf(j - box_cy(), 2);
        for (int i = 0; i < nx; ++i) {
          const float d2 = powf(i - box_cx(), 2) + y + z;
          if (d2 >= r2 && d2 <= re2) {
            const float d = sqrtf(d2);
            const float edge = (1.0 + cosf(PI * (d - mask_radius) / wanted_mask_edge)) / 2.0;
            v[address] = (R)(float)(float(v[address]) * edge + (1.0 - edge) * pixel_sum);
            mask_volume += powf(edge, 2);
          } else if (d2 >= re2) {
            v[address] = (R)(float)pixel_sum;
          } else {
            mask_volume += 1.0;
          }
          address++;
        }
        address += padding_jump_value;
      }
    }
    return float(mask_volume);
  }
  // image.cpp:6263-6340 (no mask)
  float ReturnAverageOfRealValues() const {
    double sum = 0.0;
    long address = 0;
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) { sum += v[address]; address++; }
        address += padding_jump_value;
      }
    return float(sum / (long(nx) * long(ny) * long(nz)));
  }
  // image.cpp:5907-6000
  float ReturnAverageOfRealValuesOnEdges() const {
    double sum = 0.0;
    long number_of_pixels = 0;
    long address = 0;
    if (nz == 1) {
      for (int p = 0; p < nx; ++p) { sum += v[address]; address++; }
      number_of_pixels += nx;
      address += padding_jump_value;
      for (int l = 1; l < ny - 1; ++l) {
        sum += v[address];
        address += nx - 1;
        sum += v[address];
        address += padding_jump_value + 1;
        number_of_pixels += 2;
      }
      for (int p = 0; p < nx; ++p) { sum += v[address]; address++; }
      number_of_pixels += nx;
    } else {
      for (int l = 0; l < ny; ++l) {
        for (int p = 0; p < nx; ++p) { sum += v[address]; address++; }
        address += padding_jump_value;
      }
      number_of_pixels += long(nx) * ny;
      for (int pl = 1; pl < nz - 1; ++pl) {
        for (int l = 0; l < ny; ++l) {
          if (l == 0 || l == ny - 1) {
            for (int p = 0; p < nx; ++p) { sum += v[address]; address++; }
            address += padding_jump_value;
            number_of_pixels += nx;
          } else {
            sum += v[address];
            address += nx - 1;
            sum += v[address];
            address += padding_jump_value + 1;
            number_of_pixels += 2;
          }
        }
      }
      for (int l = 0; l < ny; ++l) {
        for (int p = 0; p < nx; ++p) { sum += v[address]; address++; }
        address += padding_jump_value;
      }
      number_of_pixels += long(nx) * ny;
    }
    return sum / float(number_of_pixels);
  }
  // image.cpp:1009-1100 (no mask)
  float ReturnVarianceOfRealValues() const {
    double s = 0.0, s2 = 0.0;
    long pc = 0;
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) {
          s += v[pc];
          s2 += powf((float)v[pc], 2);
          pc++;
        }
        pc += padding_jump_value;
      }
    long n = long(nx) * long(ny) * long(nz);
    return fabsf(float(s2 / n - powf(s / n, 2)));
  }
  // image.cpp:986-1007
  void ReplaceOutliersWithMean(float maximum_n_sigmas) {
    float sigma = sqrtf(ReturnVarianceOfRealValues());
    float mean = ReturnAverageOfRealValues();
    float mx = mean + maximum_n_sigmas * sigma;
    float mn = mean - maximum_n_sigmas * sigma;
    for (long a = 0; a < real_memory_allocated; ++a) {
      if (v[a] > mx) v[a] = mean;
      else if (v[a] < mn) v[a] = mean;
    }
  }

  // --- clipping / resizing (image.cpp:8398-8464, 8522-8700, 8825-8838, 11104-11141) ---
  void ClipIntoLargerRealSpace2D(Image* o, float pad = 0.0f) const {
    o->is_in_real_space = is_in_real_space;
    o->object_is_centred_in_box = object_is_centred_in_box;
    long as = 0, ao = 0;
    const int ilo = o->box_cx() - box_cx();
    const int jlo = o->box_cy() - box_cy();
    const int ihi = ilo + nx - 1;
    const int jhi = jlo + ny - 1;
    for (int j = 0; j < o->ny; ++j) {
      if (j < jlo || j > jhi) {
        for (int i = 0; i < o->nx; ++i) { o->v[ao] = pad; ao++; }
      } else {
        for (int i = 0; i < o->nx; ++i) {
          if (i < ilo || i > ihi) o->v[ao] = pad;
          else { o->v[ao] = v[as]; as++; }
          ao++;
        }
      }
      ao += o->padding_jump_value;
      if (j >= jlo) as += padding_jump_value;
    }
  }
  // image.cpp:8522-8700 (centred, no noise fill)
  void ClipInto(Image* o, float pad = 0.0f) const {
    o->is_in_real_space = is_in_real_space;
    o->object_is_centred_in_box = object_is_centred_in_box;
    long pc = 0;
    if (is_in_real_space) {
      for (int kk = 0; kk < o->nz; ++kk) {
        int k = box_cz() + (kk - o->box_cz());
        for (int jj = 0; jj < o->ny; ++jj) {
          int j = box_cy() + (jj - o->box_cy());
          for (int ii = 0; ii < o->nx; ++ii) {
            int i = box_cx() + (ii - o->box_cx());
            if (k < 0 || k >= nz || j < 0 || j >= ny || i < 0 || i >= nx) o->v[pc] = pad;
            else o->v[pc] = v[real_addr(i, j, k)];
            pc++;
          }
          pc += o->padding_jump_value;
        }
      }
      return;
    }
    // Fourier space: other's physical (ii, jj, kk) takes this image's value at the same logical
    // coordinate (x = ii >= 0), or the padding value outside this image's logical bounds
    // (ReturnComplexPixelFromLogicalCoord, image.h:274-282)
    C* od = o->c();
    const C* d = c();
    for (int kk = 0; kk <= o->pubz(); ++kk) {
      const int lz = o->logical_z(kk);
      for (int jj = 0; jj <= o->puby(); ++jj) {
        const int ly = o->logical_y(jj);
        for (int ii = 0; ii <= o->pubx(); ++ii) {
          const int lx = ii;
          if (lx < lcx_lo() || lx > lcx_hi() || ly < lcy_lo() || ly > lcy_hi() || lz < lcz_lo() || lz > lcz_hi())
            od[pc] = C((R)pad, (R)0);
          else
            od[pc] = d[four_addr_logical(lx, ly, lz)];
          pc++;
        }
      }
    }
    // growing in y or z: the positive Nyquist planes FFTW leaves implicit are filled in
    if (ny < o->ny || nz < o->nz) {
      if (nz == 1) {
        const int jj = first_neg_y();
        for (int ii = 0; ii <= pubx(); ++ii) od[o->four_addr_phys(ii, jj, 0)] = d[four_addr_phys(ii, jj, 0)];
      } else {
        for (int kl = lcz_lo(); kl <= lcz_hi(); ++kl) {
          const int jj = first_neg_y();  // used as a logical y in the larger image, as the reference does
          const int jl = lcy_lo();
          for (int ii = 0; ii <= pubx(); ++ii) od[o->four_addr_logical(ii, jj, kl)] = d[four_addr_logical(ii, jl, kl)];
        }
        const int kk = first_neg_z();
        const int km = o->nz - first_neg_z();
        for (int jj = 1; jj <= first_neg_y(); ++jj)
          for (int ii = 0; ii <= pubx(); ++ii) od[o->four_addr_phys(ii, jj, kk)] = od[o->four_addr_phys(ii, jj, km)];
        for (int jj = o->ny - first_neg_y(); jj <= o->ny - 1; ++jj)
          for (int ii = 0; ii <= pubx(); ++ii) od[o->four_addr_phys(ii, jj, kk)] = od[o->four_addr_phys(ii, jj, km)];
        for (int ii = 0; ii <= pubx(); ++ii) od[o->four_addr_phys(ii, 0, kk)] = od[o->four_addr_phys(ii, 0, km)];
      }
    }
  }
  void Resize(int x, int y, int z, float pad) {
    if (nx == x && ny == y && nz == z) return;
    Image t;
    t.Allocate(x, y, z, is_in_real_space);
    ClipInto(&t, pad);
    *this = std::move(t);
  }
  // image.cpp:8705-8792: resample by wanted_factor (new pixel size / old) through a Fourier crop or
  // pad of an intermediately padded copy; the result is clipped into other's size.
  void ChangePixelSize(Image* o, float wanted_factor, float wanted_tolerance, bool return_fft) const {
    if (!is_in_real_space) throw std::runtime_error("oracle ChangePixelSize: image must be in real space");
    if (fabsf(wanted_factor - 1.0f) < wanted_tolerance && return_fft) {
      if (o == this) throw std::runtime_error("oracle ChangePixelSize: in place not restated");
      ClipInto(o);
      o->ForwardFFT();
      return;
    }
    if (fabsf(wanted_factor - 1.0f) < wanted_tolerance) {
      if (o != this) ClipInto(o);
      return;
    }
    int old_x = nx, old_y = ny, old_z = nz, new_x, new_y, new_z = 1, pad_x, pad_y, pad_z = 1;
    int min_dimension = std::min(nx, ny);
    if (nz > 1) min_dimension = std::min(min_dimension, nz);
    float intermediate_factor;
    float padding_value = 0.0f;
    if (wanted_tolerance < 1.0f / (min_dimension * wanted_factor)) {
      intermediate_factor = 1.0f / (wanted_tolerance * min_dimension);
      pad_x = old_x * intermediate_factor;
      if (IsEven(old_x) != IsEven(pad_x)) pad_x++;
      new_x = int(pad_x / wanted_factor);
      if (IsEven(old_x) != IsEven(new_x)) new_x++;
      pad_y = old_y * intermediate_factor;
      if (IsEven(old_y) != IsEven(pad_y)) pad_y++;
      new_y = int(pad_y / wanted_factor);
      if (IsEven(old_y) != IsEven(new_y)) new_y++;
      if (nz > 1) {
        pad_z = old_z * intermediate_factor;
        if (IsEven(old_z) != IsEven(pad_z)) pad_z++;
        new_z = int(pad_z / wanted_factor);
        if (IsEven(old_z) != IsEven(new_z)) new_z++;
      }
    } else {
      intermediate_factor = wanted_factor;
      pad_x = old_x * intermediate_factor;
      if (IsEven(old_x) != IsEven(pad_x)) pad_x++;
      new_x = old_x;
      pad_y = old_y * intermediate_factor;
      if (IsEven(old_y) != IsEven(pad_y)) pad_y++;
      new_y = old_y;
      if (nz > 1) {
        pad_z = old_z * intermediate_factor;
        if (IsEven(old_z) != IsEven(pad_z)) pad_z++;
        new_z = old_z;
      }
    }
    Image t = *this;
    if (pad_x > old_x || pad_y > old_y || pad_z > old_z) padding_value = t.ReturnAverageOfRealValues(0.45f * min_dimension);
    t.Resize(pad_x, pad_y, pad_z, padding_value);
    t.ForwardFFT();
    t.Resize(new_x, new_y, new_z, 0.0f);
    if (new_x == o->nx && new_y == o->ny && new_z == o->nz) {
      *o = std::move(t);
      if (!return_fft) o->BackwardFFT();
      return;
    }
    t.BackwardFFT();
    t.ClipInto(o, padding_value);
    if (return_fft) o->ForwardFFT();
  }
  // the sizes ChangePixelSize works at: pad and new edge per axis
  void ChangePixelSizeDims(float wanted_factor, float wanted_tolerance, int* pad, int* nw) const {
    int min_dimension = std::min(nx, ny);
    if (nz > 1) min_dimension = std::min(min_dimension, nz);
    const int dims[3] = {nx, ny, nz};
    const bool small = wanted_tolerance < 1.0f / (min_dimension * wanted_factor);
    const float intermediate_factor = small ? 1.0f / (wanted_tolerance * min_dimension) : wanted_factor;
    for (int a = 0; a < 3; ++a) {
      if (a == 2 && nz == 1) { pad[a] = nw[a] = 1; continue; }
      int p = dims[a] * intermediate_factor;
      if (IsEven(dims[a]) != IsEven(p)) p++;
      int q = small ? int(p / wanted_factor) : dims[a];
      if (small && IsEven(dims[a]) != IsEven(q)) q++;
      pad[a] = p;
      nw[a] = q;
    }
  }
  void Rotate2DInPlaceBy90Degrees(bool positive) {
    Image b;
    b.Allocate(ny, nx, 1, true);
    for (int yo = 0; yo < ny; ++yo) {
      for (int xo = 0; xo < nx; ++xo) {
        int xn, yn;
        if (positive) { xn = ny - yo - 1; yn = xo; }
        else { xn = yo; yn = nx - xo - 1; }
        b.v[b.real_addr(xn, yn, 0)] = v[real_addr(xo, yo, 0)];
      }
    }
    b.object_is_centred_in_box = object_is_centred_in_box;
    *this = std::move(b);
  }

  // --- Fourier slice extraction (image.cpp:2450-2607 without resolution limit, 2899-3020) ---
  C ReturnLinearInterpolatedFourier(float x, float y, float z) const {
    C sum(0, 0);
    const C* d = c();
    if (x >= 0.0) {
      int i_start = int(floorf(x));
      int i_end = i_start + 1;
      if (i_end > lcx_hi()) return C(0, 0);
      int j_start = int(floorf(y));
      if (j_start < lcy_lo()) return C(0, 0);
      int j_end = j_start + 1;
      if (j_end > lcy_hi()) return C(0, 0);
      int k_start = int(floorf(z));
      if (k_start < lcz_lo()) return C(0, 0);
      int k_end = k_start + 1;
      if (k_end > lcz_hi()) return C(0, 0);
      for (int k = k_start; k <= k_end; ++k) {
        int pz = k >= 0 ? k : nz + k;
        long kk = long(puby() + 1) * pz;
        float z_dist = fabsf(z - float(k));
        for (int j = j_start; j <= j_end; ++j) {
          int py = j >= 0 ? j : ny + j;
          long jj = long(pubx() + 1) * (kk + py);
          float y_dist = fabsf(y - float(j));
          for (int i = i_start; i <= i_end; ++i) {
            R weight = (1.0 - fabsf(x - float(i))) * (1.0 - y_dist) * (1.0 - z_dist);
            sum = sum + d[jj + i] * weight;
          }
        }
      }
    } else {
      int i_start = int(floorf(x));
      if (i_start < lcx_lo()) return C(0, 0);
      int i_end = i_start + 1;
      int j_start = int(floorf(y));
      if (j_start < lcy_lo()) return C(0, 0);
      int j_end = j_start + 1;
      if (j_end > lcy_hi()) return C(0, 0);
      int k_start = int(floorf(z));
      if (k_start < lcz_lo()) return C(0, 0);
      int k_end = k_start + 1;
      if (k_end > lcz_hi()) return C(0, 0);
      for (int k = k_start; k <= k_end; ++k) {
        int pz = k > 0 ? nz - k : -k;
        long kk = long(puby() + 1) * pz;
        float z_dist = fabsf(z - float(k));
        for (int j = j_start; j <= j_end; ++j) {
          int py = j > 0 ? ny - j : -j;
          long jj = long(pubx() + 1) * (kk + py);
          float y_dist = fabsf(y - float(j));
          for (int i = i_start; i <= i_end; ++i) {
            R weight = (1.0 - fabsf(x - float(i))) * (1.0 - y_dist) * (1.0 - z_dist);
            sum = sum + std::conj(d[jj - i]) * weight;
          }
        }
      }
    }
    return sum;
  }
  // m: 3x3 rotation (row-major m[r][c]); out must have the same x/y size as this volume
  // (the padding==false branch of ExtractSlice; match_template extracts at the volume size).
  void ExtractSlice(Image& out, const float m[3][3]) const {
    if (out.nx != nx || out.ny != ny) throw std::runtime_error("oracle ExtractSlice: size mismatch");
    out.object_is_centred_in_box = false;
    out.is_in_real_space = false;
    C* o = out.c();
    auto rot = [&](float xi, float yi, float& xo, float& yo, float& zo) {
      const float zi = 0.0f;
      xo = m[0][0] * xi + m[0][1] * yi + m[0][2] * zi;
      yo = m[1][0] * xi + m[1][1] * yi + m[1][2] * zi;
      zo = m[2][0] * xi + m[2][1] * yi + m[2][2] * zi;
    };
    float x3, y3, z3;
    for (int j = out.lcy_lo(); j <= out.lcy_hi(); ++j) {
      float y2 = j;
      for (int i = 1; i <= out.lcx_hi(); ++i) {
        float x2 = i;
        long pc = out.four_addr_logical(i, j, 0);
        rot(x2, y2, x3, y3, z3);
        o[pc] = ReturnLinearInterpolatedFourier(x3, y3, z3);
      }
    }
    for (int j = 1; j <= out.lcy_hi(); ++j) {
      float y2 = j, x2 = 0;
      rot(x2, y2, x3, y3, z3);
      long pc = out.four_addr_logical(0, j, 0);
      o[pc] = ReturnLinearInterpolatedFourier(x3, y3, z3);
      long pc2 = out.four_addr_logical(0, -j, 0);
      o[pc2] = std::conj(o[pc]);
    }
    if (-out.lcy_lo() != out.lcy_hi()) {
      float y2 = out.lcy_lo(), x2 = 0;
      rot(x2, y2, x3, y3, z3);
      long pc = out.four_addr_logical(0, out.lcy_lo(), 0);
      o[pc] = ReturnLinearInterpolatedFourier(x3, y3, z3);
    }
    o[0] = C(0, 0);
    out.is_in_real_space = false;
  }

  // --- spectra and filters ---
  // image.cpp:3113-3160 (real CTF, no envelope)
  void CalculateCTFImage(const CTFModel<R>& ctf);
  // image.cpp:7284-7313 (resolution_limit = 1.0)
  void ApplyCurveFilter(const Curve<R>& f, float resolution_limit = 1.0f);
  // image.cpp:7135-7205 (intensities)
  void Compute1DPowerSpectrumCurve(Curve<R>* power, Curve<R>* counts) const;
  // image.cpp:5063-5068
  bool FourierComponentIsExplicitHermitianMate(int i, int j, int k) const {
    return i == 0 && (j >= first_neg_y() || k >= first_neg_z());
  }
};

// src/core/curve.{h,cpp} subset (data in float, as the reference's Curve)
template <class R>
struct Curve {
  std::vector<float> x, y;
  int n() const { return (int)x.size(); }
  // curve.cpp:247-259
  void SetupXAxis(float lo, float hi, int np) {
    x.clear(); y.clear();
    for (int c = 0; c < np; ++c) { x.push_back(c * (hi - lo) / float(np - 1) + lo); y.push_back(0.0); }
  }
  void ZeroYData() { std::fill(y.begin(), y.end(), 0.0f); }
  // curve.cpp:834-870 (stateless: result is the unique bin with x[c] <= wx < x[c+1])
  int ReturnIndexOfNearestPreviousBin(float wx) const {
    for (int c = 0; c < n() - 1; ++c)
      if (wx >= x[c] && wx < x[c + 1]) return c;
    if (wx < x[0]) return 0;
    return n() - 1;
  }
  // curve.cpp:504-560 (assume_linear_x = true)
  void AddValueAtXUsingLinearInterpolation(float wx, float value_to_add) {
    int idx;
    if (x[0] == x[1]) idx = n() - 2;
    else idx = int((wx - x[0]) / (x[1] - x[0]));
    if (idx == n() - 1) {
      y[idx] += value_to_add;
      y[idx] += 0.0f;
    } else {
      float distance = (wx - x[idx]) / (x[idx + 1] - x[idx]);
      float vm = value_to_add * (1.0 - distance);
      float vn = value_to_add * distance;
      y[idx] += vm;
      y[idx + 1] += vn;
    }
  }
  // curve.cpp:377-396
  float ReturnLinearInterpolationFromX(float wx) const {
    float r = 0.0;
    const int i = ReturnIndexOfNearestPreviousBin(wx);
    if (i == n() - 1) r = y[n() - 1];
    else {
      float distance = (wx - x[i]) / (x[i + 1] - x[i]);
      r += y[i] * (1.0 - distance);
      r += y[i + 1] * distance;
    }
    return r;
  }
  void SquareRoot() { for (auto& a : y) a = sqrtf(a); }                      // curve.cpp:488-501
  void Reciprocal() { for (auto& a : y) if (a != 0.0) a = 1.0 / a; }        // curve.cpp:1076-1090
  void MultiplyByConstant(float k) { for (auto& a : y) a *= k; }            // curve.cpp:721-729
  float ReturnMaximumValue() const {                                        // curve.cpp:403-440
    float m = -FLT_MAX;
    for (float a : y) if (a > m) m = a;
    return m;
  }
};

// src/core/ctf.{h,cpp} subset: Init (ctf.cpp:93-150), SetDefocus (:378-383), Evaluate (:445-461)
template <class R>
struct CTFModel {
  float wavelength, squared_wavelength, cubed_wavelength, spherical_aberration, amplitude_contrast;
  float defocus_1, defocus_2, astigmatism_azimuth, additional_phase_shift;
  float precomputed_amplitude_contrast_term;
  float low_resolution_contrast = 0;
  static float WavelengthGivenAccelerationVoltage(float v) {  // ctf.cpp:556-561
    return 12.2639f / sqrtf(1000.0f * v + 0.97845e-6 * powf(1000.0f * v, 2));
  }
  void Init(float kv, float cs_mm, float amp, float df1, float df2, float astig_deg, float pixel_size,
            float phase_shift_rad) {
    wavelength = WavelengthGivenAccelerationVoltage(kv) / pixel_size;
    squared_wavelength = powf(wavelength, 2);
    cubed_wavelength = powf(wavelength, 3);
    spherical_aberration = cs_mm * 10000000.0 / pixel_size;
    amplitude_contrast = amp;
    defocus_1 = df1 / pixel_size;
    defocus_2 = df2 / pixel_size;
    astigmatism_azimuth = astig_deg / 180.0 * PI;
    additional_phase_shift = phase_shift_rad;
    if (fabs(amplitude_contrast - 1.0) < 1e-3) precomputed_amplitude_contrast_term = PI / 2.0f;
    else precomputed_amplitude_contrast_term = atanf(amplitude_contrast / sqrtf(1.0 - powf(amplitude_contrast, 2)));
  }
  void SetDefocus(float d1_px, float d2_px, float astig_rad) {
    defocus_1 = d1_px; defocus_2 = d2_px; astigmatism_azimuth = astig_rad;
  }
  float DefocusGivenAzimuth(float az) const {  // ctf.cpp:538-541
    return 0.5f * (defocus_1 + defocus_2 + cosf(2.0f * (az - astigmatism_azimuth)) * (defocus_1 - defocus_2));
  }
  float PhaseShift(float g2, float az) const {  // ctf.cpp:497-500
    return PIf * wavelength * g2 * (DefocusGivenAzimuth(az) - 0.5f * squared_wavelength * g2 * spherical_aberration) +
           additional_phase_shift + precomputed_amplitude_contrast_term;
  }
  float Evaluate(float g2, float az) const {
    if (defocus_1 == 0.0f && defocus_2 == 0.0f) return -0.7;
    return -sinf(PhaseShift(g2, az));
  }
};

template <class R>
void Image<R>::CalculateCTFImage(const CTFModel<R>& ctf) {
  long pc = 0;
  C* d = c();
  for (int j = 0; j <= puby(); ++j) {
    float y2 = logical_y(j) * fvy();
    float ysq = powf(y2, 2);
    for (int i = 0; i <= pubx(); ++i) {
      float x2 = i * fvx();
      float az = (i == 0 && j == 0) ? 0.0f : atan2f(y2, x2);
      float f2 = powf(x2, 2) + ysq;
      d[pc] = C((R)ctf.Evaluate(f2, az), (R)0);
      pc++;
    }
  }
}

template <class R>
void Image<R>::ApplyCurveFilter(const Curve<R>& f, float resolution_limit) {
  long pc = 0;
  C* d = c();
  for (int k = 0; k <= pubz(); ++k) {
    float sz = std::pow(logical_z(k) * fvz(), 2);
    for (int j = 0; j <= puby(); ++j) {
      float sy = std::pow(logical_y(j) * fvy(), 2);
      for (int i = 0; i <= pubx(); ++i) {
        float sx = std::pow(i * fvx(), 2);
        float s2 = sx + sy + sz;
        float sf = std::sqrt((double)s2);
        if (sf <= resolution_limit) d[pc] *= (R)f.ReturnLinearInterpolationFromX(sf);
        else d[pc] = C(0, 0);
        pc++;
      }
    }
  }
}

template <class R>
void Image<R>::Compute1DPowerSpectrumCurve(Curve<R>* power, Curve<R>* counts) const {
  power->ZeroYData();
  counts->ZeroYData();
  long address = 0;
  const C* d = c();
  for (int k = 0; k <= pubz(); ++k) {
    float sz = powf(logical_z(k) * fvz(), 2);
    for (int j = 0; j <= puby(); ++j) {
      float sy = powf(logical_y(j) * fvy(), 2);
      for (int i = 0; i <= pubx(); ++i) {
        if (FourierComponentIsExplicitHermitianMate(i, j, k)) { address++; continue; }
        float sx = powf(i * fvx(), 2);
        float sf = sqrtf(sx + sy + sz);
        float re = (float)d[address].real(), im = (float)d[address].imag();
        power->AddValueAtXUsingLinearInterpolation(sf, re * re + im * im);
        counts->AddValueAtXUsingLinearInterpolation(sf, 1.0);
        address++;
      }
    }
  }
  for (int c2 = 0; c2 < power->n(); ++c2) {
    if (counts->y[c2] > 0.0) power->y[c2] /= counts->y[c2];
    else power->y[c2] = 0.0;
  }
}

// src/core/matrix.cpp:490-515 (ZYZ Euler matrix from degrees)
inline void EulerMatrix(float phi, float theta, float psi, float m[3][3]) {
  float cp = cosf(deg_2_rad(phi)), sp = sinf(deg_2_rad(phi));
  float ct = cosf(deg_2_rad(theta)), st = sinf(deg_2_rad(theta));
  float cs = cosf(deg_2_rad(psi)), ss = sinf(deg_2_rad(psi));
  m[0][0] = cp * ct * cs - sp * ss;
  m[1][0] = sp * ct * cs + cp * ss;
  m[2][0] = -st * cs;
  m[0][1] = -cp * ct * ss - sp * cs;
  m[1][1] = -sp * ct * ss + cp * cs;
  m[2][1] = st * ss;
  m[0][2] = st * cp;
  m[1][2] = st * sp;
  m[2][2] = ct;
}

}  // namespace tmo
