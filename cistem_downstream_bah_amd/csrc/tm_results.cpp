// Worker and master result records (include/tm_results.h): host code, no device work.
#include "../../include/tm_results.h"

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_rerr;

struct ResultError {
  int code;
  std::string msg;
};

template <class F>
int rguard(F&& f) {
  try {
    f();
    return TM_OK;
  } catch (const ResultError& e) {
    g_rerr = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    g_rerr = e.what();
    return TM_ERR_INVALID;
  }
}

void require(bool cond, const std::string& msg) {
  if (!cond) throw ResultError{TM_ERR_INVALID, msg};
}

// padding_jump_value of a real-space image of width x (2 for even x, 1 for odd x)
int row_jump(int x) { return (x % 2 == 0) ? 2 : 1; }

long long map_floats(int x, int y) { return (long long)(x + row_jump(x)) * y; }

// Image::ReturnAverageOfRealValues(radius, false) of a dense nx x ny image (image.cpp:6263-6345):
// the mean over pixels whose squared distance from (nx / 2, ny / 2), in float, is <= radius^2
// (FP64 sum), or over all pixels when radius <= 0; 0 when the disc holds no pixel.
float disc_average(const float* img, int nx, int ny, float radius) {
  double sum = 0.0;
  if (radius > 0.0f) {
    const float r2 = radius * radius;
    const int cx = nx / 2, cy = ny / 2;
    long long n = 0;
    for (int j = 0; j < ny; ++j) {
      const float dy = (float)(j - cy);
      const float y2 = dy * dy;
      for (int i = 0; i < nx; ++i) {
        const float dx = (float)(i - cx);
        if (dx * dx + y2 + 0.0f <= r2) {
          sum += img[(size_t)j * nx + i];
          ++n;
        }
      }
    }
    return n > 0 ? (float)(sum / (double)n) : 0.0f;
  }
  for (long long e = 0; e < (long long)nx * ny; ++e) sum += img[e];
  return (float)(sum / (double)((long long)nx * ny));
}

// Image::Resize to (ox, oy) with a padding value (ClipInto centred, image.cpp:8522-8575):
// output pixel (ii, jj) reads input (nx / 2 + ii - ox / 2, ny / 2 + jj - oy / 2) or the fill.
std::vector<float> resize(const std::vector<float>& in, int nx, int ny, int ox, int oy, float fill) {
  if (nx == ox && ny == oy) return in;
  std::vector<float> out((size_t)ox * oy);
  for (int jj = 0; jj < oy; ++jj) {
    const int j = ny / 2 + (jj - oy / 2);
    for (int ii = 0; ii < ox; ++ii) {
      const int i = nx / 2 + (ii - ox / 2);
      out[(size_t)jj * ox + ii] = (i < 0 || i >= nx || j < 0 || j >= ny) ? fill : in[(size_t)j * nx + i];
    }
  }
  return out;
}

struct Header {
  int x, y, bins;
  long long n;
};

Header read_header(const float* r, long long floats) {
  require(floats >= TM_RESULT_META, "record shorter than its header");
  Header h;
  h.x = (int)r[0];
  h.y = (int)r[1];
  h.n = (long long)r[2];
  h.bins = (int)r[3];
  require(h.x > 0 && h.y > 0 && h.bins >= 0 && h.n == map_floats(h.x, h.y), "record header inconsistent");
  require(floats == tm_result_size(h.x, h.y, h.bins), "record size does not match its header");
  return h;
}

}  // namespace

extern "C" {

// cisTEM_erfinv / cisTEM_erfcinv (functions.cpp:1282-1299): Winitzki's closed-form inverse of his
// approximation erf(x) ~ sqrt(1 - exp(-x^2 (4/pi + a x^2) / (1 + a x^2))), a = 0.147, in double.
static double winitzki_erfinv(double x) {
  const double PI = 3.14159265359;  // defines.h:6-7
  const double a = 0.147;
  const double sgn = (x < 0) ? -1.0 : 1.0;
  const double lnx = log((1 - x) * (1 + x));
  const double tt1 = 2 / (PI * a) + 0.5 * lnx;
  const double tt2 = 1 / a * lnx;
  return sgn * sqrt(-tt1 + sqrt(tt1 * tt1 - tt2));
}

double tm_expected_threshold(int image_x, int image_y, long long total_ccfs) {
  const double erf_input = 2.0 / (1.0 * (double)image_x * (double)image_y * (double)total_ccfs);
  return (double)(sqrtf(2.0f) * (float)winitzki_erfinv(1.0 - erf_input));
}

int tm_histogram_table(const long long* histogram, int bins, float histogram_min, float histogram_step,
                       double sqrt_input_pixels, long long total_ccfs, float* snr, double* survival,
                       double* expected) {
  return rguard([&] {
    require(histogram && bins > 0, "histogram is NULL or empty");
    const float first = histogram_min + (histogram_step / 2.0f);
    const float scale = (float)(sqrt_input_pixels * sqrt_input_pixels) * float(total_ccfs);
    double run = 0.0;
    for (int i = bins - 1; i >= 0; --i) {
      run += (double)histogram[i];
      if (survival) survival[i] = run;
    }
    for (int i = 0; i < bins; ++i) {
      const float x = first + histogram_step * float(i);
      if (snr) snr[i] = x;
      if (expected) expected[i] = (std::erfc(x / sqrtf(2.0f)) / 2.0f) * scale;
    }
  });
}

const char* tm_result_last_error(void) { return g_rerr.c_str(); }

long long tm_result_size(int x, int y, int histogram_bins) {
  if (x <= 0 || y <= 0 || histogram_bins < 0) return -1;
  return TM_RESULT_META + map_floats(x, y) * TM_RESULT_MAPS + 2LL * histogram_bins;
}

int tm_result_pack(const float* const* maps, int nx, int ny, int out_x, int out_y, int remove_npix_from_edge,
                   const long long* histogram, int histogram_bins, float n_ccfs, double sqrt_input_pixels, float* out,
                   long long out_floats) {
  return rguard([&] {
    require(maps && out && (histogram || histogram_bins == 0), "NULL argument");
    for (int k = 0; k < TM_RESULT_MAPS; ++k) require(maps[k] != nullptr, "NULL map");
    require(nx > 0 && ny > 0 && out_x > 0 && out_y > 0 && histogram_bins >= 0, "sizes must be positive");
    require(remove_npix_from_edge >= 0, "remove_npix_from_edge must be >= 0");
    require(out_floats == tm_result_size(out_x, out_y, histogram_bins), "out_floats != tm_result_size(out_x, out_y, bins)");
    // match_template.cpp:1393-1406
    int trim_x = out_x - remove_npix_from_edge, trim_y = out_y - remove_npix_from_edge;
    if (out_x > nx) trim_x = nx - remove_npix_from_edge;
    if (out_y > ny) trim_y = ny - remove_npix_from_edge;
    require(trim_x > 0 && trim_y > 0, "remove_npix_from_edge leaves no valid area");
    const float central_region = (float)(0.35 * (double)(nx + ny - 2 * remove_npix_from_edge));
    const long long n = map_floats(out_x, out_y);
    const int jump = row_jump(out_x);
    out[0] = (float)out_x;
    out[1] = (float)out_y;
    out[2] = (float)n;
    out[3] = (float)histogram_bins;
    out[4] = n_ccfs;
    out[5] = (float)sqrt_input_pixels;
    for (int k = 0; k < TM_RESULT_MAPS; ++k) {
      std::vector<float> m(maps[k], maps[k] + (size_t)nx * ny);
      const float avg = disc_average(m.data(), nx, ny, central_region);
      m = resize(m, nx, ny, trim_x, trim_y, avg);
      m = resize(m, trim_x, trim_y, out_x, out_y, avg);
      float* dst = out + TM_RESULT_META + k * n;
      for (int j = 0; j < out_y; ++j) {
        float* row = dst + (size_t)j * (out_x + jump);
        std::memcpy(row, &m[(size_t)j * out_x], out_x * sizeof(float));
        for (int p = 0; p < jump; ++p) row[out_x + p] = 0.0f;
      }
    }
    if (histogram_bins)
      std::memcpy(out + TM_RESULT_META + TM_RESULT_MAPS * n, histogram, (size_t)histogram_bins * sizeof(long long));
  });
}

int tm_result_add(float* collated, const float* result, long long floats, float* total_ccfs, int first) {
  return rguard([&] {
    require(collated && result && total_ccfs, "NULL argument");
    const Header h = read_header(result, floats);
    if (first) {
      std::memset(collated, 0, (size_t)floats * sizeof(float));
      for (int i = 0; i < 4; ++i) collated[i] = result[i];
      collated[5] = result[5];
      *total_ccfs = 0.0f;
    } else {
      const Header c = read_header(collated, floats);
      require(c.x == h.x && c.y == h.y && c.bins == h.bins, "records describe different maps");
    }
    *total_ccfs += result[4];
    const long long n = h.n;
    float* cm = collated + TM_RESULT_META;
    const float* rm = result + TM_RESULT_META;
    for (long long p = 0; p < n; ++p) {
      if (rm[p] > cm[p])
        for (int k = 0; k < 6; ++k) cm[k * n + p] = rm[k * n + p];
    }
    for (long long p = 0; p < n; ++p) cm[6 * n + p] += rm[6 * n + p];
    for (long long p = 0; p < n; ++p) cm[7 * n + p] += rm[7 * n + p];
    long long ch, rh;
    for (int b = 0; b < h.bins; ++b) {
      std::memcpy(&ch, cm + TM_RESULT_MAPS * n + 2 * b, sizeof(ch));
      std::memcpy(&rh, rm + TM_RESULT_MAPS * n + 2 * b, sizeof(rh));
      ch += rh;
      std::memcpy(cm + TM_RESULT_MAPS * n + 2 * b, &ch, sizeof(ch));
    }
  });
}

int tm_result_finalize(const float* collated, long long floats, float total_ccfs, float* mip_scaled, float* zscore,
                       float* average, float* std_dev, float* angles, long long* histogram) {
  return rguard([&] {
    require(collated != nullptr, "NULL argument");
    const Header h = read_header(collated, floats);
    require(total_ccfs > 0.0f, "total_ccfs must be positive");
    const long long n = h.n;
    const int jump = row_jump(h.x);
    const double sqrt_px = collated[5];
    const float* cm = collated + TM_RESULT_META;
    for (int j = 0; j < h.y; ++j)
      for (int i = 0; i < h.x; ++i) {
        const long long p = (long long)j * (h.x + jump) + i;
        const size_t o = (size_t)j * h.x + i;
        const float mip = cm[p];
        const float mean = cm[6 * n + p] / total_ccfs;
        float sd = sqrtf(cm[7 * n + p] / total_ccfs - powf(mean, 2.0f));
        float z;
        if (sd > 0.0f) {
          z = (mip - mean) / sd;
        } else {
          sd = 0.0f;
          z = 0.0f;
        }
        if (mip_scaled) mip_scaled[o] = (float)(mip * sqrt_px);
        if (zscore) zscore[o] = z;
        if (average) average[o] = (float)(mean * sqrt_px);
        if (std_dev) std_dev[o] = (float)(sd * sqrt_px);
        if (angles)
          for (int k = 0; k < 5; ++k) angles[(size_t)k * h.x * h.y + o] = cm[(k + 1) * n + p];
      }
    if (histogram && h.bins)
      std::memcpy(histogram, cm + TM_RESULT_MAPS * n, (size_t)h.bins * sizeof(long long));
  });
}

}  // extern "C"
