// TEST INFRASTRUCTURE ONLY — part of the CPU oracle for the template-matching path.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this code,
// and only as the checker; the product never links or calls it.
//
// Mixed-radix Stockham FFT (radices 4, 2 and a generic odd-prime DFT of any size) with
// FFTW3's conventions for the reference's in-place "padded" real layout:
//   * R2C output of a row of nx reals is nx/2+1 complex values in the same row
//     (row stride nx+2 floats), src/core/image.cpp:4806-4829 (fftwf_plan_dft_r2c_2d/3d).
//   * Forward unscaled here; Image::ForwardFFT scales by 1/N afterwards
//     (src/core/image.cpp:5085-5102). Inverse C2R is unscaled (image.cpp:5104-5116).
// FFTW itself is a third-party dependency of the reference whose version is not pinned
// (configure.ac:289-320); this file restates the DFT, not FFTW's operation order.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <vector>

namespace tmo {

template <class R>
struct FFTPlan {
  int n = 1;
  std::vector<int> radix;               // factors, applied in order
  std::vector<std::complex<R>> tw;      // tw[k] = exp(-2 pi i k / n), k < n
  std::vector<std::complex<R>> half_tw; // exp(-2 pi i k / (2n)), k <= n   (real-FFT post-processing)
  std::map<int, std::vector<std::complex<R>>> roots;  // roots[r][j] = exp(-2 pi i j / r)
};

template <class R>
inline std::shared_ptr<const FFTPlan<R>> fft_plan(int n) {
  static std::mutex mu;
  static std::map<int, std::shared_ptr<const FFTPlan<R>>> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(n);
  if (it != cache.end()) return it->second;
  auto p = std::make_shared<FFTPlan<R>>();
  p->n = n;
  int m = n;
  while (m % 4 == 0) { p->radix.push_back(4); m /= 4; }
  if (m % 2 == 0) { p->radix.push_back(2); m /= 2; }
  for (int f = 3; m > 1; f += 2) {
    while (m % f == 0) { p->radix.push_back(f); m /= f; }
  }
  const long double two_pi = 2.0L * std::acos(-1.0L);
  p->tw.resize(n);
  for (int k = 0; k < n; ++k) {
    long double a = two_pi * (long double)k / (long double)n;
    p->tw[k] = std::complex<R>((R)std::cos(a), (R)-std::sin(a));
  }
  p->half_tw.resize(n + 1);
  for (int k = 0; k <= n; ++k) {
    long double a = two_pi * (long double)k / (long double)(2 * n);
    p->half_tw[k] = std::complex<R>((R)std::cos(a), (R)-std::sin(a));
  }
  for (int r : p->radix) {
    if (p->roots.count(r)) continue;
    std::vector<std::complex<R>> rt(r);
    for (int j = 0; j < r; ++j) {
      long double a = two_pi * (long double)j / (long double)r;
      rt[j] = std::complex<R>((R)std::cos(a), (R)-std::sin(a));
    }
    p->roots[r] = rt;
  }
  cache[n] = p;
  return p;
}

// In-register DFT of size r; sign = -1 forward, +1 inverse.
template <class R>
inline void small_dft(std::complex<R>* v, int r, int sign, const std::complex<R>* rt) {
  typedef std::complex<R> C;
  if (r == 2) {
    C a = v[0], b = v[1];
    v[0] = a + b; v[1] = a - b;
    return;
  }
  if (r == 4) {
    C t0 = v[0] + v[2], t1 = v[0] - v[2], t2 = v[1] + v[3];
    C d = v[1] - v[3];
    C t3 = (sign < 0) ? C(d.imag(), -d.real()) : C(-d.imag(), d.real());
    v[0] = t0 + t2; v[2] = t0 - t2; v[1] = t1 + t3; v[3] = t1 - t3;
    return;
  }
  // generic odd radix: O(r^2) with the plan's exact roots
  C small[16];
  std::vector<C> big(r > 16 ? r : 0);
  C* out = r > 16 ? big.data() : small;
  for (int k = 0; k < r; ++k) {
    C acc(0, 0);
    for (int q = 0; q < r; ++q) {
      C w = rt[(q * k) % r];
      if (sign > 0) w = std::conj(w);
      acc += v[q] * w;
    }
    out[k] = acc;
  }
  for (int k = 0; k < r; ++k) v[k] = out[k];
}

// Complex FFT of length p.n, in place on x, using scratch y (length n). Unscaled.
template <class R>
inline void fft_c2c(const FFTPlan<R>& p, std::complex<R>* x, std::complex<R>* y, int sign) {
  typedef std::complex<R> C;
  const int n = p.n;
  if (n == 1) return;
  C* src = x;
  C* dst = y;
  int Ns = 1;
  int rmax = 16;
  for (int r : p.radix) rmax = std::max(rmax, r);
  std::vector<C> vbuf(rmax);
  C* v = vbuf.data();
  for (int r : p.radix) {
    const C* rt = p.roots.at(r).data();
    const int m = n / r;
    const int step = n / (Ns * r);
    const int groups = m / Ns;
    for (int g = 0; g < groups; ++g) {
      for (int k = 0; k < Ns; ++k) {
        const int j = g * Ns + k;
        for (int q = 0; q < r; ++q) v[q] = src[j + q * m];
        if (Ns > 1) {
          for (int q = 1; q < r; ++q) {
            C w = p.tw[(long)k * q * step];
            if (sign > 0) w = std::conj(w);
            v[q] *= w;
          }
        }
        small_dft(v, r, sign, rt);
        const int base = g * Ns * r + k;
        for (int q = 0; q < r; ++q) dst[base + q * Ns] = v[q];
      }
    }
    std::swap(src, dst);
    Ns *= r;
  }
  if (src != x) std::memcpy((void*)x, (const void*)src, sizeof(C) * n);
}

// Real-to-complex of one row of n reals (n even) stored in place: floats [0, n) in,
// n/2+1 complex out over floats [0, n+2). Unscaled.
template <class R>
inline void rfft_row(R* row, int n, std::complex<R>* scratch) {
  typedef std::complex<R> C;
  const int h = n / 2;
  auto p = fft_plan<R>(h);
  C* z = reinterpret_cast<C*>(row);
  fft_c2c(*p, z, scratch, -1);
  // post-process pairs (k, h-k); W^k = exp(-2 pi i k / n) = half_tw[k] of the h-point plan
  C z0 = z[0];
  C X0(z0.real() + z0.imag(), 0);
  C Xh(z0.real() - z0.imag(), 0);
  for (int k = 1; k <= h / 2; ++k) {
    const int kk = h - k;
    C zk = z[k], zc = std::conj(z[kk]);
    C E = (zk + zc) * (R)0.5;
    C D = (zk - zc) * (R)0.5;
    C O(D.imag(), -D.real());  // D / i
    C w = p->half_tw[k];
    C Xk = E + w * O;
    C Xkk = std::conj(E - w * O);
    z[k] = Xk;
    z[kk] = Xkk;
  }
  z[0] = X0;
  z[h] = Xh;
}

// Complex-to-real inverse (unscaled, FFTW c2r semantics: imaginary parts of the DC and
// Nyquist terms are ignored). n/2+1 complex in over floats [0, n+2), n reals out.
template <class R>
inline void irfft_row(R* row, int n, std::complex<R>* scratch) {
  typedef std::complex<R> C;
  const int h = n / 2;
  auto p = fft_plan<R>(h);
  C* z = reinterpret_cast<C*>(row);
  const R x0 = z[0].real(), xh = z[h].real();
  for (int k = 1; k <= h / 2; ++k) {
    const int kk = h - k;
    C Xk = z[k], Xc = std::conj(z[kk]);
    C E = Xk + Xc;
    C w = std::conj(p->half_tw[k]);
    C O = (Xk - Xc) * w;
    C Zk = E + C(-O.imag(), O.real());       // E + i*O
    // partner h-k: E' = conj(E), O' = conj(Xk - Xc) * W^k  (since conj(W^(h-k)) = -W^k)
    C Ek = std::conj(E);
    C Ok = std::conj(Xk - Xc) * p->half_tw[k];
    C Zkk = Ek + C(-Ok.imag(), Ok.real());
    z[k] = Zk;
    z[kk] = Zkk;
  }
  z[0] = C(x0 + xh, x0 - xh);
  fft_c2c(*p, z, scratch, +1);
}

}  // namespace tmo
