// A caller shaped like match_template's GPU branch (src/programs/match_template/match_template.cpp:
// 858-985), compiled against include/template_matching_core.hpp with stand-in types that carry the
// reference's member names (Image, EulerSearch, AnglesAndShifts, ProgressBar, MyApp). It calls the
// reference-signature Init and RunInnerLoop, then merges each plane on the host from
// d_max_intensity_projection / d_best_* / d_sum3 / d_sumSq3 / histogram exactly as that loop does:
// strict-greater MIP with the plane's angles and defocus offset, double sums of the float plane
// sums, histogram counts added. Test infrastructure: tests/test_gpu_dropin.py writes the inputs and
// compares the result with the C ABI's search.
//
//   reference_caller DIR   (reads DIR/meta.txt and DIR/*.bin, writes DIR/out.bin)
#include <cfloat>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "template_matching_core.hpp"

// Stand-ins with the member names the reference's caller and TemplateMatchingCore::Init use.
struct MyApp {};
struct ProgressBar {};
struct AnglesAndShifts {};
struct EulerSearch {
  int number_of_search_positions = 0;
  float** list_of_search_parameters = nullptr;
};
struct Image {
  int logical_x_dimension = 0, logical_y_dimension = 0, logical_z_dimension = 1, padding_jump_value = 2;
  std::vector<float> storage;
  float* real_values = nullptr;
  std::complex<float>* complex_values = nullptr;
  void Allocate(int x, int y, int z) {
    logical_x_dimension = x;
    logical_y_dimension = y;
    logical_z_dimension = z;
    padding_jump_value = x % 2 == 0 ? 2 : 1;
    storage.assign((size_t)(x + padding_jump_value) * y * z, 0.f);
    real_values = storage.data();
    complex_values = reinterpret_cast<std::complex<float>*>(storage.data());
  }
  void SetToConstant(float v) {
    for (auto& a : storage) a = v;
  }
};

static std::vector<float> read_floats(const std::string& path, size_t n) {
  std::vector<float> v(n);
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f || std::fread(v.data(), sizeof(float), n, f) != n) {
    std::fprintf(stderr, "cannot read %s\n", path.c_str());
    std::exit(2);
  }
  std::fclose(f);
  return v;
}

int main(int argc, char** argv) {
  if (argc != 2) return 2;
  const std::string dir = argv[1];
  int P, W, H, T, n_pos, first, last, nbins, n_def_half;
  float psi_start, psi_max, psi_step, defocus_step, hist_min, hist_step;
  FILE* m = std::fopen((dir + "/meta.txt").c_str(), "r");
  if (!m || std::fscanf(m, "%d %d %d %d %d %d %d %d %d %f %f %f %f %f %f", &P, &W, &H, &T, &n_pos, &first, &last,
                        &nbins, &n_def_half, &psi_start, &psi_max, &psi_step, &defocus_step, &hist_min,
                        &hist_step) != 15) {
    std::fprintf(stderr, "cannot parse %s/meta.txt\n", dir.c_str());
    return 2;
  }
  std::fclose(m);
  Image template_reconstruction, input_image, current_projection;
  template_reconstruction.Allocate(P, P, P);
  auto tv = read_floats(dir + "/template.bin", (size_t)P * P * (P + 2));
  std::copy(tv.begin(), tv.end(), template_reconstruction.storage.begin());
  input_image.Allocate(W, H, 1);
  auto iv = read_floats(dir + "/input.bin", (size_t)H * (W + 2));
  std::copy(iv.begin(), iv.end(), input_image.storage.begin());
  current_projection.Allocate(T, T, 1);
  auto gv = read_floats(dir + "/grid.bin", 2 * (size_t)n_pos);
  std::vector<float*> rows(n_pos);
  for (int i = 0; i < n_pos; ++i) rows[i] = &gv[2 * (size_t)i];
  EulerSearch global_euler_search;
  global_euler_search.number_of_search_positions = n_pos;
  global_euler_search.list_of_search_parameters = rows.data();
  AnglesAndShifts angles;
  MyApp app;
  ProgressBar* my_progress = nullptr;
  const int n_planes = 2 * n_def_half + 1;
  auto fv = read_floats(dir + "/filters.bin", (size_t)n_planes * T * (T + 2));

  // host maps as the caller holds them (match_template.cpp:600-632)
  Image max_intensity_projection, best_psi, best_theta, best_phi, best_defocus, best_pixel_size;
  for (Image* im : {&max_intensity_projection, &best_psi, &best_theta, &best_phi, &best_defocus, &best_pixel_size})
    im->Allocate(W, H, 1);
  max_intensity_projection.SetToConstant(-FLT_MAX);
  const size_t numel = (size_t)(W + 2) * H;
  std::vector<double> correlation_pixel_sum(numel, 0.0), correlation_pixel_sum_of_squares(numel, 0.0);
  std::vector<long> histogram_data(nbins, 0);
  long actual_number_of_ccs_calculated = 0, current_correlation_position = 0;
  const float pixel_size_step = 0.02f;

  cistem_amd::TemplateMatchingCore* GPU = new cistem_amd::TemplateMatchingCore[1];
  const int tIDX = 0;
  GPU[tIDX].Init(&app, template_reconstruction, input_image, current_projection, 0.0f, pixel_size_step, 1.0f,
                 defocus_step * n_def_half, defocus_step, 15000.0f, 15000.0f, psi_max, psi_start, psi_step, angles,
                 global_euler_search, hist_min, hist_step, nbins, 0, first, last, my_progress,
                 (long)(last - first + 1) * 1, true);
  const int size_i = 0;
  for (int defocus_i = -n_def_half; defocus_i <= n_def_half; ++defocus_i) {
    Image projection_filter;
    projection_filter.Allocate(T, T, 1);
    const float* f = &fv[(size_t)(defocus_i + n_def_half) * T * (T + 2)];
    std::copy(f, f + (size_t)T * (T + 2), projection_filter.storage.begin());
    GPU[tIDX].RunInnerLoop(projection_filter, size_i, defocus_i, tIDX, current_correlation_position);
    // the merge of match_template.cpp:918-985
    Image mip_buffer, psi_buffer, phi_buffer, theta_buffer, sum, sumSq;
    for (Image* im : {&mip_buffer, &psi_buffer, &phi_buffer, &theta_buffer, &sum, &sumSq}) im->Allocate(W, H, 1);
    GPU[tIDX].d_max_intensity_projection.CopyDeviceToHost(mip_buffer, true, false);
    GPU[tIDX].d_best_psi.CopyDeviceToHost(psi_buffer, true, false);
    GPU[tIDX].d_best_phi.CopyDeviceToHost(phi_buffer, true, false);
    GPU[tIDX].d_best_theta.CopyDeviceToHost(theta_buffer, true, false);
    GPU[tIDX].d_sum3.CopyDeviceToHost(sum, true, false);
    GPU[tIDX].d_sumSq3.CopyDeviceToHost(sumSq, true, false);
    GPU[tIDX].d_max_intensity_projection.Wait();
    size_t pixel_counter = 0;
    for (int current_y = 0; current_y < H; ++current_y) {
      for (int current_x = 0; current_x < W; ++current_x) {
        if (mip_buffer.real_values[pixel_counter] > max_intensity_projection.real_values[pixel_counter]) {
          max_intensity_projection.real_values[pixel_counter] = mip_buffer.real_values[pixel_counter];
          best_psi.real_values[pixel_counter] = psi_buffer.real_values[pixel_counter];
          best_theta.real_values[pixel_counter] = theta_buffer.real_values[pixel_counter];
          best_phi.real_values[pixel_counter] = phi_buffer.real_values[pixel_counter];
          best_defocus.real_values[pixel_counter] = float(defocus_i) * defocus_step;
          best_pixel_size.real_values[pixel_counter] = float(size_i) * pixel_size_step;
        }
        correlation_pixel_sum[pixel_counter] += (double)sum.real_values[pixel_counter];
        correlation_pixel_sum_of_squares[pixel_counter] += (double)sumSq.real_values[pixel_counter];
        pixel_counter++;
      }
      pixel_counter += max_intensity_projection.padding_jump_value;
    }
    GPU[tIDX].histogram.CopyToHostAndAdd(histogram_data.data());
    actual_number_of_ccs_calculated += GPU[tIDX].total_number_of_cccs_calculated;
  }
  delete[] GPU;

  FILE* o = std::fopen((dir + "/out.bin").c_str(), "wb");
  for (Image* im : {&max_intensity_projection, &best_psi, &best_theta, &best_phi, &best_defocus})
    std::fwrite(im->real_values, sizeof(float), numel, o);
  std::fwrite(correlation_pixel_sum.data(), sizeof(double), numel, o);
  std::fwrite(correlation_pixel_sum_of_squares.data(), sizeof(double), numel, o);
  std::vector<long long> h64(histogram_data.begin(), histogram_data.end());
  std::fwrite(h64.data(), sizeof(long long), h64.size(), o);
  const long long counts[2] = {actual_number_of_ccs_calculated, current_correlation_position};
  std::fwrite(counts, sizeof(long long), 2, o);
  std::fclose(o);
  return 0;
}
