// C++ drop-in for the reference's GPU class TemplateMatchingCore
// (src/gpu/TemplateMatchingCore.h:55-179), header-only over the C ABI of tm_hip.h.
//
// The reference passes cisTEM `Image` objects; this class takes the same data as plain arrays in
// the same memory layouts (FFTW half-complex spectra, padded real maps), so a caller holding
// Images passes image.real_values / reinterpret_cast<float*>(image.complex_values). Method names
// and argument meaning follow the reference:
//   Init(...)          TemplateMatchingCore.h:147-171 (search ranges, Euler grid, psi loop, histogram)
//   RunInnerLoop(...)  TemplateMatchingCore.h:173 (one pixel-size/defocus plane; size_i and
//                      defocus_i are loop indices, as in the reference)
// The accumulated maps (the reference's d_max_intensity_projection, d_best_psi/theta/phi,
// d_sum3/d_sumSq3 and histogram, read by match_template.cpp:918-985) are fetched with
// CopyResults(). The merge over planes happens on the device in the reference's loop order, so
// the caller does not merge per plane. Errors throw std::runtime_error (the reference exits).
#ifndef TEMPLATE_MATCHING_CORE_HPP
#define TEMPLATE_MATCHING_CORE_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "tm_hip.h"

namespace cistem_amd {

class TemplateMatchingCore {
 public:
  explicit TemplateMatchingCore(int device = 0) {
    check(tm_create(device, &h_));
  }
  ~TemplateMatchingCore() { tm_destroy(h_); }
  TemplateMatchingCore(const TemplateMatchingCore&) = delete;
  TemplateMatchingCore& operator=(const TemplateMatchingCore&) = delete;

  // template_reconstruction: P x P x (P/2+1) complex (quadrant-swapped, DC 0)
  // input_image:             H x (W/2+1) complex (whitened, normalised, quadrant-swapped)
  // projection_size:         edge T of current_projection
  // psi_max/psi_start/psi_step: the inclusive float loop of match_template.cpp:1001
  // search_positions:        global_euler_search.list_of_search_parameters as {phi, theta} pairs
  void Init(const float* template_reconstruction, int template_size, const float* input_image, int image_x,
            int image_y, int projection_size, float pixel_size_step, float defocus_step, float psi_max,
            float psi_start, float psi_step, const float* search_positions, int number_of_search_positions,
            double histogram_min_scaled, double histogram_step_scaled, int histogram_number_of_bins,
            int max_padding, int first_search_position, int last_search_position, int batch_size = 0,
            int plane_storage = TM_PLANE_FP32) {
    psi_.clear();
    for (float p = psi_start; p <= psi_max; p += psi_step) psi_.push_back(p);
    tm_init_params prm{};
    prm.template_size = template_size;
    prm.projection_size = projection_size;
    prm.image_x = image_x;
    prm.image_y = image_y;
    prm.n_search_positions = number_of_search_positions;
    prm.n_psi = (int)psi_.size();
    prm.first_search_position = first_search_position;
    prm.last_search_position = last_search_position;
    prm.histogram_min_scaled = histogram_min_scaled;
    prm.histogram_step_scaled = histogram_step_scaled;
    prm.histogram_number_of_bins = histogram_number_of_bins;
    prm.max_padding = max_padding;
    prm.defocus_step = defocus_step;
    prm.pixel_size_step = pixel_size_step;
    prm.batch_size = batch_size;
    prm.plane_storage = plane_storage;
    check(tm_init(h_, &prm, template_reconstruction, input_image, search_positions, psi_.data()));
    image_x_ = image_x;
    image_y_ = image_y;
    n_bins_ = histogram_number_of_bins;
  }

  // projection_filter: T x (T/2+1) complex (CTF x whitening). Returns the CCFs computed.
  long long RunInnerLoop(const float* projection_filter, int pixel_i, int defocus_i,
                         tm_progress_fn progress = nullptr, void* user = nullptr) {
    long long n = 0;
    check(tm_run_inner_loop(h_, projection_filter, pixel_i, defocus_i, progress, user, &n));
    total_number_of_cccs_calculated += n;
    return n;
  }

  // Padded-layout maps ((W+2) * H); any pointer may be null. histogram is added to (like
  // Histogram::CopyToHostAndAdd, Histogram.cu:188-203).
  void CopyResults(float* mip, float* psi, float* theta, float* phi, float* defocus, float* pixel_size, double* sum,
                   double* sum_of_squares, long* histogram_to_add) {
    std::vector<long long> hist(512, 0);
    check(tm_get_results(h_, mip, psi, theta, phi, defocus, pixel_size, sum, sum_of_squares, nullptr, hist.data(),
                         nullptr));
    if (histogram_to_add)
      for (int i = 0; i < n_bins_; ++i) histogram_to_add[i] += (long)hist[i];
  }

  // The next pixel-size plane's template (same size and layout as at Init); the maps keep
  // accumulating, as after the reference's GPU[tIDX].template_reconstruction.CopyFrom
  // (match_template.cpp:883-886).
  void SetTemplate(const float* template_spectrum) { check(tm_set_template(h_, template_spectrum)); }
  // Prepares a micrograph at the search size on the device into the input spectrum, resets the maps
  // (no counterpart in the reference).
  void SetMicrograph(const float* micrograph, float* curve_x, float* curve_y) {
    check(tm_set_micrograph(h_, micrograph, curve_x, curve_y));
  }
  // Replaces the input spectrum and resets the maps (no counterpart in the reference).
  void SetInput(const float* input_spectrum) { check(tm_set_input(h_, input_spectrum)); }
  // Peaks of the z-score map on the device, in the layout of TemplateMatchFoundPeakInfo
  // (template_matching.h:4-18); returns their number (no counterpart in the reference's GPU path).
  int FindPeaks(long long total_ccfs, float threshold, int edge, float min_peak_radius, float pixel_size,
                tm_peak_info* out, int capacity) {
    int n = 0;
    check(tm_find_peaks(h_, total_ccfs, threshold, edge, min_peak_radius, pixel_size, out, capacity, &n));
    return n;
  }
  // Replays each plane from one captured hipGraph (no counterpart in the reference).
  void SetGraph(bool enabled) { check(tm_set_graph(h_, enabled ? 1 : 0)); }

  // Folds another core's results into this one (same search, another position range, same device),
  // as AggregatedTemplateResult::AddResult merges job results (match_template.cpp:2015-2103).
  void Merge(TemplateMatchingCore& other) {
    check(tm_merge(h_, other.h_));
    total_number_of_cccs_calculated += other.total_number_of_cccs_calculated;
  }

  // Final scaling and z-score on the device (match_template.cpp:1221-1243, 1300-1310); padded
  // layout outputs, any pointer may be null.
  void Finalize(long long total_correlation_positions, float* mip_scaled, float* zscore, float* average,
                float* std_dev) {
    check(tm_finalize(h_, total_correlation_positions, mip_scaled, zscore, average, std_dev));
  }

  long long total_number_of_cccs_calculated = 0;

 private:
  static void check(int rc) {
    if (rc != TM_OK) throw std::runtime_error(std::string("TemplateMatchingCore: ") + tm_last_error());
  }
  tm_core* h_ = nullptr;
  std::vector<float> psi_;
  int image_x_ = 0, image_y_ = 0, n_bins_ = 512;
};

}  // namespace cistem_amd

#endif  // TEMPLATE_MATCHING_CORE_HPP
