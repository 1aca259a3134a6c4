// Device-side FFT building blocks for gfx950 (CDNA4, 64-wide wavefronts).
//
// Stockham autosort formulation: for a length-N transform built from radices R_1..R_s, stage s
// with current sub-transform length Ns has N/R butterflies; butterfly j reads
// x[j + q*N/R] (q < R), multiplies by W^(k*q*N/(Ns*R)) with k = j mod Ns, does a DFT-R in
// registers and writes y[(j/Ns)*Ns*R + k + q*Ns]. Twiddles come from a per-N table of
// exp(-2 pi i k / N) computed on the host in long double (exact to float rounding).
//
// Conventions follow FFTW, which the reference uses (src/core/image.cpp:4806-4829,
// 5085-5116): forward sign -1 unscaled, inverse sign +1 unscaled; the reference's 1/N forward
// scaling is applied by the callers.
#pragma once

#include <hip/hip_runtime.h>

namespace tmk {

struct cf { float x, y; };

__device__ __forceinline__ cf cmake(float a, float b) { cf r; r.x = a; r.y = b; return r; }
__device__ __forceinline__ cf cadd(cf a, cf b) { return cmake(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cf csub(cf a, cf b) { return cmake(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cf cmul(cf a, cf b) { return cmake(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ cf cconj(cf a) { return cmake(a.x, -a.y); }
__device__ __forceinline__ cf cscale(cf a, float s) { return cmake(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <int SIGN>
__device__ __forceinline__ cf cmul_mi(cf a) { return SIGN < 0 ? cmake(a.y, -a.x) : cmake(-a.y, a.x); }
// conj(a) * b : the reference's conj(reference) * input (match_template.cpp:1083)
__device__ __forceinline__ cf cconjmul(cf a, cf b) { return cmake(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x); }

// ---------------------------------------------------------------- register DFTs
template <int SIGN>
__device__ __forceinline__ void dft2(cf& a, cf& b) { cf t = a; a = cadd(t, b); b = csub(t, b); }

template <int SIGN>
__device__ __forceinline__ void dft4(cf* v) {
  cf t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  cf t2 = cadd(v[1], v[3]), t3 = cmul_mi<SIGN>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2); v[2] = csub(t0, t2); v[1] = cadd(t1, t3); v[3] = csub(t1, t3);
}

// radix-3: X1,2 = x0 - (x1 + x2)/2 +- SIGN i sin(2 pi/3) (x1 - x2)
template <int SIGN>
__device__ __forceinline__ void dft3(cf* v) {
  const float s3 = 0.86602540378443864676f;
  const cf t1 = cadd(v[1], v[2]);
  const cf t2 = cmake(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
  const cf u = cscale(cmul_mi<SIGN>(csub(v[1], v[2])), s3);
  v[0] = cadd(v[0], t1);
  v[1] = cadd(t2, u);
  v[2] = csub(t2, u);
}

// radix-5 with c_k = cos(2 pi k/5), s_k = sin(2 pi k/5), a = x1 + x4, a' = x2 + x3, b = x1 - x4,
// b' = x2 - x3: X1,4 = x0 + c1 a + c2 a' +- SIGN i (s1 b + s2 b'), X2,3 = x0 + c2 a + c1 a' +- SIGN i (s2 b - s1 b')
template <int SIGN>
__device__ __forceinline__ void dft5(cf* v) {
  const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
  const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
  const cf a1 = cadd(v[1], v[4]), a2 = cadd(v[2], v[3]);
  const cf b1 = csub(v[1], v[4]), b2 = csub(v[2], v[3]);
  const cf p1 = cmake(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
  const cf p2 = cmake(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
  const cf q1 = cmul_mi<SIGN>(cmake(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
  const cf q2 = cmul_mi<SIGN>(cmake(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
  v[0] = cadd(v[0], cadd(a1, a2));
  v[1] = cadd(p1, q1);
  v[4] = csub(p1, q1);
  v[2] = cadd(p2, q2);
  v[3] = csub(p2, q2);
}

// radix-8: 2 x 4 with internal twiddles W8^k
template <int SIGN>
__device__ __forceinline__ void dft8(cf* v) {
  const float r = 0.70710678118654752440f;
  cf a[4] = {v[0], v[2], v[4], v[6]};
  cf b[4] = {v[1], v[3], v[5], v[7]};
  dft4<SIGN>(a);
  dft4<SIGN>(b);
  // b[k] *= W8^(k), W8 = exp(SIGN * 2 pi i / 8)
  b[1] = cmake((b[1].x - SIGN * b[1].y) * r, (b[1].y + SIGN * b[1].x) * r);  // * (1 + SIGN i)/sqrt2
  b[2] = cmul_mi<SIGN>(b[2]);
  b[3] = cmake((-b[3].x - SIGN * b[3].y) * r, (SIGN * b[3].x - b[3].y) * r);  // * (-1 + SIGN i)/sqrt2
  for (int k = 0; k < 4; ++k) { v[k] = cadd(a[k], b[k]); v[k + 4] = csub(a[k], b[k]); }
}

// radix-16: 4 x 4 with internal twiddles W16^(k1*k2)
template <int SIGN>
__device__ __forceinline__ void dft16(cf* v) {
  // column DFT-4 over stride-4 subsequences
  cf t[4][4];
#pragma unroll
  for (int n1 = 0; n1 < 4; ++n1) {
    cf s[4] = {v[n1], v[n1 + 4], v[n1 + 8], v[n1 + 12]};
    dft4<SIGN>(s);
#pragma unroll
    for (int k = 0; k < 4; ++k) t[n1][k] = s[k];
  }
  // twiddles W16^(n1*k2) with W16 = exp(SIGN 2 pi i/16)
  const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f, c2 = 0.70710678118654752440f;
  // n1=1: k=1 (c1, s1), k=2 (c2, c2), k=3 (s1, c1)
  // n1=2: k=1 (c2, c2), k=2 (0, 1), k=3 (-c2, c2)
  // n1=3: k=1 (s1, c1), k=2 (-c2, c2), k=3 (-c1, -s1)
  auto tw = [](cf a, float c, float s) { return cmake(a.x * c - a.y * (SIGN * s), a.x * (SIGN * s) + a.y * c); };
  t[1][1] = tw(t[1][1], c1, s1);
  t[1][2] = tw(t[1][2], c2, c2);
  t[1][3] = tw(t[1][3], s1, c1);
  t[2][1] = tw(t[2][1], c2, c2);
  t[2][2] = cmake(-SIGN * t[2][2].y, SIGN * t[2][2].x);
  t[2][3] = tw(t[2][3], -c2, c2);
  t[3][1] = tw(t[3][1], s1, c1);
  t[3][2] = tw(t[3][2], -c2, c2);
  t[3][3] = tw(t[3][3], -c1, -s1);
  // row DFT-4 over n1 for each k1; output index k1 + 4*k2
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    cf s[4] = {t[0][k1], t[1][k1], t[2][k1], t[3][k1]};
    dft4<SIGN>(s);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = s[k2];
  }
}

// radix-9: 3 x 3 with internal twiddles W9^(n1*k1); input n = n1 + 3 m, output k1 + 3 k2
template <int SIGN>
__device__ __forceinline__ void dft9(cf* v) {
  cf t[3][3];
#pragma unroll
  for (int n1 = 0; n1 < 3; ++n1) {
    cf s[3] = {v[n1], v[n1 + 3], v[n1 + 6]};
    dft3<SIGN>(s);
#pragma unroll
    for (int k = 0; k < 3; ++k) t[n1][k] = s[k];
  }
  // W9^m = cos(2 pi m/9) + SIGN i sin(2 pi m/9) for m = n1 * k1 = 1, 2, 4
  const float c1 = 0.76604444311897803520f, s1 = 0.64278760968653932632f;
  const float c2 = 0.17364817766693034885f, s2 = 0.98480775301220805936f;
  const float c4 = -0.93969262078590838405f, s4 = 0.34202014332566873304f;
  auto tw = [](cf a, float c, float s) { return cmake(a.x * c - a.y * (SIGN * s), a.x * (SIGN * s) + a.y * c); };
  t[1][1] = tw(t[1][1], c1, s1);
  t[1][2] = tw(t[1][2], c2, s2);
  t[2][1] = tw(t[2][1], c2, s2);
  t[2][2] = tw(t[2][2], c4, s4);
#pragma unroll
  for (int k1 = 0; k1 < 3; ++k1) {
    cf s[3] = {t[0][k1], t[1][k1], t[2][k1]};
    dft3<SIGN>(s);
#pragma unroll
    for (int k2 = 0; k2 < 3; ++k2) v[k1 + 3 * k2] = s[k2];
  }
}

template <int R, int SIGN>
__device__ __forceinline__ void dftR(cf* v) {
  if constexpr (R == 2) dft2<SIGN>(v[0], v[1]);
  else if constexpr (R == 4) dft4<SIGN>(v);
  else if constexpr (R == 8) dft8<SIGN>(v);
  else if constexpr (R == 16) dft16<SIGN>(v);
  else if constexpr (R == 3) dft3<SIGN>(v);
  else if constexpr (R == 9) dft9<SIGN>(v);
  else if constexpr (R == 5) dft5<SIGN>(v);
  else static_assert(R == 2 || R == 3 || R == 4 || R == 5 || R == 8 || R == 9 || R == 16, "unsupported radix");
}

// Workgroup barrier that orders LDS traffic only: this wave's LDS operations complete, then
// s_barrier. __syncthreads() also waits for every outstanding global load and store of the wave
// (vmcnt 0), which would stall on Z stores or prefetches that the barrier does not need to order.
// The memory clobber keeps the compiler from moving memory accesses across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Barrier after direct global-to-LDS loads: they count as vector-memory operations (vmcnt), so
// each wave waits for its own before the barrier makes all of them visible to the workgroup.
__device__ __forceinline__ void lds_dma_barrier() {
  asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// LDS atomic add without return (ds_add_u32). Written out because the compiler's wait
// insertion cannot tell an atomic apart from a pending direct global-to-LDS load and would wait for
// every such load (vmcnt 0) before each histogram update. The instruction has no output register;
// LDS operations complete in order, so the compiler's own LDS waits stay correct, and an
// lds_barrier() completes it before the counts are read.
__device__ __forceinline__ void lds_add(unsigned int* p, unsigned int v) {
  const unsigned addr = (unsigned)(size_t)((__attribute__((address_space(3))) unsigned int*)p);
  asm volatile("ds_add_u32 %0, %1" ::"v"(addr), "v"(v) : "memory");
}

// twiddle W^(idx) from the table (forward table); conjugated for the inverse.
template <int SIGN>
__device__ __forceinline__ cf twiddle(const cf* __restrict__ tw, int idx) {
  cf w = tw[idx];
  return SIGN < 0 ? w : cconj(w);
}

// Factored twiddle table: W^m = lo[m mod 64] * hi[m / 64], with lo[b] = W^b and hi[a] = W^(64 a).
// Two small tables replace one of length N (for tables kept in LDS).
struct TwSplit {
  const cf* lo;
  const cf* hi;
};
template <int SIGN>
__device__ __forceinline__ cf twiddle(const TwSplit& t, int idx) {
  const cf a = t.lo[idx & 63], b = t.hi[idx >> 6];
  const cf w = cmake(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
  return SIGN < 0 ? w : cconj(w);
}

// The twiddle W^(k q N / (NS R)) of butterfly input q at sub-transform index k of a Stockham stage.
template <int SIGN, int N, int NS, int R, typename TW>
__device__ __forceinline__ cf stage_twiddle(TW tw, int k, int q) {
  return twiddle<SIGN>(tw, k * q * (N / (NS * R)));
}
// Factored table plus a direct table of the 64th roots of unity (d64[m] = W^(m N / 64)): the stages
// whose angles are all multiples of 2 pi / 64 (NS R <= 64) read their twiddles directly.
struct TwHybrid {
  const cf* d64;
  TwSplit split;
};
template <int SIGN, int N, int NS, int R>
__device__ __forceinline__ cf stage_twiddle(const TwHybrid& tw, int k, int q) {
  if constexpr (NS * R <= 64 && 64 % (NS * R) == 0) {
    const cf w = tw.d64[k * q * (64 / (NS * R))];
    return SIGN < 0 ? w : cconj(w);
  } else {
    return twiddle<SIGN>(tw.split, k * q * (N / (NS * R)));
  }
}

// All twiddles w[q] = W^(k q N / (NS R)), q = 1 .. R-1, of one butterfly (w[0] unused).
template <int SIGN, int N, int NS, int R, typename TW>
__device__ __forceinline__ void stage_twiddles(TW tw, int k, cf* w) {
#pragma unroll
  for (int q = 1; q < R; ++q) w[q] = stage_twiddle<SIGN, N, NS, R>(tw, k, q);
}
// A full twiddle table in global memory, read once per butterfly: w[1] = W^(k N / (NS R)) is looked
// up and the other twiddles of the butterfly are its powers (w[2q] = w[q]^2, w[2q+1] = w[2q] w[1]).
struct TwPow {
  const cf* t;
};
template <int SIGN, int N, int NS, int R>
__device__ __forceinline__ void stage_twiddles(const TwPow& tw, int k, cf* w) {
  w[1] = twiddle<SIGN>(tw.t, k * (N / (NS * R)));
#pragma unroll
  for (int q = 2; q < R; ++q) w[q] = (q & 1) ? cmul(w[q - 1], w[1]) : cmul(w[q / 2], w[q / 2]);
}

// TwHybrid: stages past the 64th roots look up w[1] only and form the other powers by products
// (w[2q] = w[q]^2, w[2q+1] = w[2q] w[1]: at most three products deep for R <= 8).
template <int SIGN, int N, int NS, int R>
__device__ __forceinline__ void stage_twiddles(const TwHybrid& tw, int k, cf* w) {
  if constexpr (NS * R <= 64 && 64 % (NS * R) == 0) {
#pragma unroll
    for (int q = 1; q < R; ++q) w[q] = stage_twiddle<SIGN, N, NS, R>(tw, k, q);
  } else {
    w[1] = stage_twiddle<SIGN, N, NS, R>(tw, k, 1);
#pragma unroll
    for (int q = 2; q < R; ++q) w[q] = (q & 1) ? cmul(w[q - 1], w[1]) : cmul(w[q / 2], w[q / 2]);
  }
}

// ---------------------------------------------------------------- radix plans
// N = 2^a 3^b 5^c. The power-of-two part runs first, largest radix first: radices 2^LR (LR = 4:
// radix 16, LR = 3: radix 8) and one smaller radix when a is not a multiple of LR; then the factors
// of three as radix-9 stages (3 x 3 in registers) and a radix-3 stage when b is odd; then c radix-5
// stages. For powers of two this is the plain largest-radix-first plan.
__host__ __device__ constexpr int log2i(int n) { return n <= 1 ? 0 : 1 + log2i(n / 2); }
__host__ __device__ constexpr int nfactor(int n, int f) { return (n > 0 && n % f == 0) ? 1 + nfactor(n / f, f) : 0; }
__host__ __device__ constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }
__host__ __device__ constexpr bool fft_size_ok(int n) {
  return n >= 1 && ipow(2, nfactor(n, 2)) * ipow(3, nfactor(n, 3)) * ipow(5, nfactor(n, 5)) == n;
}
__host__ __device__ constexpr int plan_stages2(int N, int LR) {
  return nfactor(N, 2) / LR + ((nfactor(N, 2) % LR) ? 1 : 0);
}
__host__ __device__ constexpr int plan_stages3(int N) { return nfactor(N, 3) / 2 + nfactor(N, 3) % 2; }
__host__ __device__ constexpr int plan_stages(int N, int LR = 4) {
  return plan_stages2(N, LR) + plan_stages3(N) + nfactor(N, 5);
}
__host__ __device__ constexpr int plan_radix(int N, int s, int LR = 4) {
  return s < plan_stages2(N, LR)
             ? ((s < nfactor(N, 2) / LR) ? (1 << LR) : (1 << (nfactor(N, 2) % LR)))
             : (s - plan_stages2(N, LR) < nfactor(N, 3) / 2 ? 9 : (s - plan_stages2(N, LR) < plan_stages3(N) ? 3 : 5));
}
__host__ __device__ constexpr int plan_ns(int N, int s, int LR = 4) {  // sub-transform length before stage s
  int ns = 1;
  for (int i = 0; i < s; ++i) ns *= plan_radix(N, i, LR);
  return ns;
}

// One Stockham stage executed by `nthreads` threads over LDS buffers (in != out allowed to alias
// only when the caller separates loads and stores with a barrier; here in/out are distinct).
template <int N, int R, int NS, int SIGN>
__device__ __forceinline__ void stage_lds(const cf* __restrict__ in, cf* __restrict__ out, const cf* __restrict__ tw,
                                          int tid, int nthreads) {
  constexpr int M = N / R;
  for (int j = tid; j < M; j += nthreads) {
    cf v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = in[j + q * M];
    const int k = j % NS;
    if constexpr (NS > 1) {
      constexpr int step = N / (NS * R);
#pragma unroll
      for (int q = 1; q < R; ++q) v[q] = cmul(v[q], twiddle<SIGN>(tw, k * q * step));
    }
    dftR<R, SIGN>(v);
    const int base = (j / NS) * NS * R + k;
#pragma unroll
    for (int q = 0; q < R; ++q) out[base + q * NS] = v[q];
  }
}

// Full power-of-two FFT of one line held in LDS buffer a (scratch b, both N long).
// Result ends in a if the number of stages is even, else in b; returns the result pointer.
template <int N, int SIGN, int S = 0>
__device__ __forceinline__ cf* fft_lds(cf* a, cf* b, const cf* __restrict__ tw, int tid, int nthreads) {
  if constexpr (S >= plan_stages(N)) {
    return a;
  } else {
    stage_lds<N, plan_radix(N, S), plan_ns(N, S), SIGN>(a, b, tw, tid, nthreads);
    __syncthreads();
    return fft_lds<N, SIGN, S + 1>(b, a, tw, tid, nthreads);
  }
}

}  // namespace tmk
