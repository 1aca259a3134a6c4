// Host-side harness for the result-record code (cistem_downstream_bah_amd/csrc/tm_results.cpp, the
// worker/master records of include/tm_results.h), built with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_sanitize_cpu.py. It packs worker records of odd and even
// map sizes with edge cuts and padding, folds them into a collated record, finalises it, evaluates
// the expected threshold and the histogram table, and drives every error path (NULL pointers, bad
// sizes, mismatched records); any out-of-bounds access, use of uninitialised heap, signed overflow or
// misaligned load stops the run with a sanitizer report. Exit status 0 when every check holds.
#include <cmath>
#include <cstdio>
#include <vector>

#include "tm_results.h"

static int failures = 0;
#define CHECK(c)                                                      \
  do {                                                                \
    if (!(c)) {                                                       \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      ++failures;                                                     \
    }                                                                 \
  } while (0)

static std::vector<float> record(int nx, int ny, int out_x, int out_y, int remove, int bins, float base) {
  std::vector<std::vector<float>> maps(8, std::vector<float>((size_t)nx * ny));
  for (int k = 0; k < 8; ++k)
    for (int i = 0; i < nx * ny; ++i) maps[k][i] = base + 0.001f * (float)((i * 7 + k * 13) % 101);
  const float* mp[8];
  for (int k = 0; k < 8; ++k) mp[k] = maps[k].data();
  std::vector<long long> hist(bins);
  for (int b = 0; b < bins; ++b) hist[b] = (long long)(b % 5) * 3;
  const long long n = tm_result_size(out_x, out_y, bins);
  CHECK(n > 0);
  std::vector<float> out((size_t)n, -1.f);
  CHECK(tm_result_pack(mp, nx, ny, out_x, out_y, remove, hist.data(), bins, 1000.0f, std::sqrt((float)(nx * ny)),
                       out.data(), n) == TM_OK);
  return out;
}

int main() {
  const int bins = 512;
  for (int shape = 0; shape < 4; ++shape) {
    const int nx = shape & 1 ? 37 : 40, ny = shape & 2 ? 29 : 32;
    const int out_x = nx + (shape == 3 ? 6 : 0), out_y = ny + (shape == 1 ? 3 : 0);
    const int remove = shape;
    std::vector<float> a = record(nx, ny, out_x, out_y, remove, bins, 0.5f);
    std::vector<float> b = record(nx, ny, out_x, out_y, remove, bins, 0.6f);
    const long long n = (long long)a.size();
    std::vector<float> col((size_t)n, 0.f);
    float total = -1.f;
    CHECK(tm_result_add(col.data(), a.data(), n, &total, 1) == TM_OK);
    CHECK(tm_result_add(col.data(), b.data(), n, &total, 0) == TM_OK);
    CHECK(total == 2000.0f);
    const size_t px = (size_t)out_x * out_y;
    std::vector<float> mip(px), z(px), avg(px), sd(px), ang(5 * px);
    std::vector<long long> hist(bins);
    CHECK(tm_result_finalize(col.data(), n, total, mip.data(), z.data(), avg.data(), sd.data(), ang.data(),
                             hist.data()) == TM_OK);
    CHECK(hist[1] == 6 && hist[5] == 0);
    for (size_t i = 0; i < px; ++i) CHECK(std::isfinite(z[i]) && std::isfinite(mip[i]));
    CHECK(tm_result_finalize(col.data(), n, total, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr) == TM_OK);
    // mismatched records and bad arguments are errors, not crashes
    std::vector<float> small = record(nx, ny, nx, ny, 0, 16, 0.5f);
    CHECK(tm_result_add(col.data(), small.data(), (long long)small.size(), &total, 0) != TM_OK);
    CHECK(tm_result_add(nullptr, a.data(), n, &total, 0) != TM_OK);
    CHECK(tm_result_finalize(col.data(), n - 1, total, mip.data(), nullptr, nullptr, nullptr, nullptr, nullptr) !=
          TM_OK);
    CHECK(tm_result_last_error() != nullptr);
  }
  const float* none[8] = {nullptr};
  std::vector<long long> h(4, 1);
  std::vector<float> out(16);
  CHECK(tm_result_pack(none, 4, 4, 4, 4, 0, h.data(), 4, 1.f, 4.f, out.data(), 16) != TM_OK);
  CHECK(tm_result_size(0, 4, 4) <= 0 || tm_result_size(-1, 4, 4) <= 0);
  // expected threshold and histogram table
  const double thr = tm_expected_threshold(4096, 4096, 2200000);
  CHECK(thr > 7.0 && thr < 8.5);
  std::vector<long long> counts(bins);
  for (int i = 0; i < bins; ++i) counts[i] = (long long)((i * 37) % 11);
  std::vector<float> snr(bins);
  std::vector<double> surv(bins), expct(bins);
  CHECK(tm_histogram_table(counts.data(), bins, -12.5f, 35.0f / 512, 4096.0, 1000, snr.data(), surv.data(),
                           expct.data()) == TM_OK);
  long long tail = 0;
  for (int i = bins - 1; i >= 0; --i) {
    tail += counts[i];
    CHECK(surv[i] == (double)tail);
  }
  CHECK(tm_histogram_table(nullptr, bins, -12.5f, 0.07f, 4096.0, 1000, snr.data(), nullptr, nullptr) != TM_OK);
  if (failures) {
    std::fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  std::printf("result-record harness: all checks passed\n");
  return 0;
}
