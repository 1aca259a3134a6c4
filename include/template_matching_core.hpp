// C++ drop-in for the reference's GPU class TemplateMatchingCore
// (src/gpu/TemplateMatchingCore.h:55-179), header-only over the C ABI of tm_hip.h.
//
// Two forms. The reference's own signatures, templated on the caller's Image / EulerSearch types
// (Init with the 25 arguments of TemplateMatchingCore.h:147-171, RunInnerLoop(Image&, float, float,
// int, long&) of :173, and the members d_max_intensity_projection, d_best_psi/theta/phi, d_sum3,
// d_sumSq3, histogram and total_number_of_cccs_calculated that match_template.cpp:918-985 reads), so
// that caller compiles unchanged and merges planes on the host as it does now. And a plain-array
// form in the same memory layouts (FFTW half-complex spectra, padded real maps) whose maps accumulate
// over planes on the device. Method names and argument meaning follow the reference:
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

  // ------------------------------------------------------------------------------------------
  // The reference's own signatures (TemplateMatchingCore.h:147-173), templated on the caller's
  // types so that match_template.cpp:858-985 compiles against this class unchanged. ImageT is the
  // reference's Image: complex_values (std::complex<float>*, FFTW half-complex), real_values,
  // logical_x/y_dimension, padding_jump_value. EulerT provides number_of_search_positions and
  // list_of_search_parameters ({phi, theta} rows). parent_pointer, angles, my_progress,
  // total_correlation_positions, is_running_locally and threadIDX carry the reference's progress and
  // job plumbing and are not used; the template, input and current_projection sizes give P, W x H and T.
  template <class AppT, class ImageT, class AnglesT, class EulerT, class ProgressT>
  void Init(AppT* parent_pointer, ImageT& template_reconstruction, ImageT& input_image, ImageT& current_projection,
            float pixel_size_search_range, float pixel_size_step, float pixel_size, float defocus_search_range,
            float defocus_step, float defocus1, float defocus2, float psi_max, float psi_start, float psi_step,
            AnglesT& angles, EulerT& global_euler_search, float histogram_min_scaled, float histogram_step_scaled,
            int histogram_number_of_bins, int max_padding, int first_search_position, int last_search_position,
            ProgressT* my_progress, long total_correlation_positions, bool is_running_locally) {
    (void)parent_pointer; (void)pixel_size_search_range; (void)pixel_size; (void)defocus_search_range;
    (void)defocus1; (void)defocus2; (void)angles; (void)my_progress; (void)total_correlation_positions;
    (void)is_running_locally;
    const int n = global_euler_search.number_of_search_positions;
    std::vector<float> grid(2 * (size_t)n);
    for (int i = 0; i < n; ++i) {
      grid[2 * (size_t)i] = global_euler_search.list_of_search_parameters[i][0];
      grid[2 * (size_t)i + 1] = global_euler_search.list_of_search_parameters[i][1];
    }
    Init(reinterpret_cast<const float*>(template_reconstruction.complex_values),
         template_reconstruction.logical_x_dimension, reinterpret_cast<const float*>(input_image.complex_values),
         input_image.logical_x_dimension, input_image.logical_y_dimension, current_projection.logical_x_dimension,
         pixel_size_step, defocus_step, psi_max, psi_start, psi_step, grid.data(), n, histogram_min_scaled,
         histogram_step_scaled, histogram_number_of_bins, max_padding, first_search_position, last_search_position);
  }

  // One plane, as the reference's RunInnerLoop (TemplateMatchingCore.cu:140-368): the maps, sums and
  // histogram start from zero (-FLT_MAX for the MIP, the CPU path's start) at every call, so the
  // caller merges planes on the host (match_template.cpp:918-985) from the members below;
  // total_number_of_cccs_calculated is this call's count and current_correlation_position grows by it.
  template <class ImageT>
  void RunInnerLoop(ImageT& projection_filter, float pixel_i, float defocus_i, int threadIDX,
                    long& current_correlation_position) {
    (void)threadIDX;
    check(tm_reset(h_));
    cache_valid_ = false;
    long long n = 0;
    check(tm_run_inner_loop(h_, reinterpret_cast<const float*>(projection_filter.complex_values), (int)pixel_i,
                            (int)defocus_i, nullptr, nullptr, &n));
    total_number_of_cccs_calculated = n;
    current_correlation_position += (long)n;
  }

  // Stand-ins for the reference's GpuImage members that match_template.cpp:922-947 reads:
  // CopyDeviceToHost(Image& dst, bool block, bool free) fills dst's padded real-space rows; Wait() is
  // a no-op (copies complete before returning).
  class DeviceMap {
   public:
    DeviceMap(TemplateMatchingCore* owner, int which) : owner_(owner), which_(which) {}
    template <class ImageT>
    void CopyDeviceToHost(ImageT& dst, bool should_block_until_complete = true, bool free_gpu_memory = false) {
      (void)should_block_until_complete; (void)free_gpu_memory;
      owner_->copy_map(which_, dst.real_values, dst.logical_x_dimension, dst.logical_y_dimension,
                       dst.padding_jump_value);
    }
    void Wait() {}

   private:
    TemplateMatchingCore* owner_;
    int which_;
  };
  // Stand-in for the reference's Histogram member (Histogram.cu:188-203).
  class HistogramCopy {
   public:
    explicit HistogramCopy(TemplateMatchingCore* owner) : owner_(owner) {}
    void CopyToHostAndAdd(long* histogram_data) { owner_->add_histogram(histogram_data); }

   private:
    TemplateMatchingCore* owner_;
  };
  DeviceMap d_max_intensity_projection{this, 0}, d_best_psi{this, 1}, d_best_theta{this, 2}, d_best_phi{this, 3},
      d_sum3{this, 4}, d_sumSq3{this, 5};
  HistogramCopy histogram{this};

  long long total_number_of_cccs_calculated = 0;

 private:
  static void check(int rc) {
    if (rc != TM_OK) throw std::runtime_error(std::string("TemplateMatchingCore: ") + tm_last_error());
  }
  void fetch() {
    if (cache_valid_) return;
    const size_t n = (size_t)(image_x_ + 2) * image_y_;
    for (auto& m : maps_) m.assign(n, 0.f);
    sums_[0].assign(n, 0.0);
    sums_[1].assign(n, 0.0);
    hist_.assign(512, 0);
    check(tm_get_results(h_, maps_[0].data(), maps_[1].data(), maps_[2].data(), maps_[3].data(), nullptr, nullptr,
                         sums_[0].data(), sums_[1].data(), nullptr, hist_.data(), nullptr));
    cache_valid_ = true;
  }
  // map 0 MIP, 1 psi, 2 theta, 3 phi, 4 sum, 5 sum of squares (as floats) into rows of x + jump floats
  void copy_map(int which, float* dst, int x, int y, int jump) {
    if (x != image_x_ || y != image_y_) throw std::runtime_error("TemplateMatchingCore: image size differs");
    fetch();
    for (int j = 0; j < y; ++j)
      for (int i = 0; i < x; ++i) {
        const size_t s = (size_t)j * (image_x_ + 2) + i;
        dst[(size_t)j * (x + jump) + i] = which < 4 ? maps_[which][s] : (float)sums_[which - 4][s];
      }
  }
  void add_histogram(long* histogram_data) {
    fetch();
    for (int b = 0; b < n_bins_; ++b) histogram_data[b] += (long)hist_[b];
  }
  tm_core* h_ = nullptr;
  std::vector<float> psi_;
  int image_x_ = 0, image_y_ = 0, n_bins_ = 512;
  bool cache_valid_ = false;
  std::vector<float> maps_[4];
  std::vector<double> sums_[2];
  std::vector<long long> hist_;
};

}  // namespace cistem_amd

#endif  // TEMPLATE_MATCHING_CORE_HPP
