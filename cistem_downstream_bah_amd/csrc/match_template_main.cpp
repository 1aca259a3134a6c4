// match_template-style command-line driver (C++ host) over libtm_hip.so.
//
// Mirrors MatchTemplateApp::DoCalculation for a local run on one GPU
// (src/programs/match_template/match_template.cpp:283-1370): prepare the image and template,
// build the Euler grid and in-plane angles, loop over defocus planes through TemplateMatchingCore,
// then scale, z-score, un-rotate, crop to the original size and write the maps as MRC files,
// the histogram text file (:1333-1364, master form :1763-1837) and the peaks above the expected
// threshold (:1839-1980, iterative maximum with radius exclusion).
//
//   match_template_hip --image img.mrc --template vol.mrc --out prefix [options]
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cctype>
#include <fstream>
#include <iostream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/template_matching_core.hpp"
#include "../../include/tm_prepare.h"
#include "../../include/tm_results.h"

namespace {

struct Vol {
  int nx = 0, ny = 0, nz = 0;
  std::vector<float> v;  // x fastest
};

Vol read_mrc(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  int32_t hdr[256];
  f.read(reinterpret_cast<char*>(hdr), 1024);
  if (hdr[3] != 2) throw std::runtime_error(path + ": only MRC mode 2 is supported");
  Vol r;
  r.nx = hdr[0]; r.ny = hdr[1]; r.nz = hdr[2];
  f.seekg(1024 + hdr[23]);
  r.v.resize((size_t)r.nx * r.ny * r.nz);
  f.read(reinterpret_cast<char*>(r.v.data()), r.v.size() * sizeof(float));
  if (!f) throw std::runtime_error(path + ": truncated data");
  return r;
}

void write_mrc(const std::string& path, const float* data, int nx, int ny, float pixel_size) {
  int32_t hdr[256] = {0};
  hdr[0] = nx; hdr[1] = ny; hdr[2] = 1; hdr[3] = 2;
  hdr[7] = nx; hdr[8] = ny; hdr[9] = 1;
  float cell[3] = {nx * pixel_size, ny * pixel_size, pixel_size};
  std::memcpy(&hdr[10], cell, sizeof(cell));
  float ang[3] = {90.f, 90.f, 90.f};
  std::memcpy(&hdr[13], ang, sizeof(ang));
  hdr[16] = 1; hdr[17] = 2; hdr[18] = 3;
  float mn = FLT_MAX, mx = -FLT_MAX;
  double mean = 0;
  for (long i = 0; i < (long)nx * ny; ++i) { mn = std::min(mn, data[i]); mx = std::max(mx, data[i]); mean += data[i]; }
  float stats[3] = {mn, mx, (float)(mean / ((double)nx * ny))};
  std::memcpy(&hdr[19], stats, sizeof(stats));
  std::memcpy(&hdr[52], "MAP ", 4);
  hdr[53] = 0x00004444;
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(hdr), 1024);
  f.write(reinterpret_cast<const char*>(data), (size_t)nx * ny * sizeof(float));
}

// Padded-layout W x H map -> dense and un-rotated (Rotate2DInPlaceBy90Degrees(false),
// image.cpp:11127-11138); nx, ny receive the dense size.
template <class T>
std::vector<float> dense_unrotated(const T* padded, int W, int H, bool rotated, int* nx, int* ny) {
  std::vector<float> d((size_t)W * H);
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) d[(size_t)y * W + x] = (float)padded[(size_t)y * (W + 2) + x];
  *nx = W;
  *ny = H;
  if (rotated) {
    std::vector<float> r((size_t)W * H);
    const int rx = H;  // rotated image is H x W
    for (int yo = 0; yo < H; ++yo)
      for (int xo = 0; xo < W; ++xo) r[(size_t)(W - xo - 1) * rx + yo] = d[(size_t)yo * W + xo];
    d.swap(r);
    *nx = H;
    *ny = W;
  }
  return d;
}

// Padded-layout W x H map -> dense, un-rotated and cropped/padded to the original size with a fill
// value (Image::Resize -> ClipInto, image.cpp:8522-8575, 8825-8838).
std::vector<float> to_original(const float* padded, int W, int H, bool rotated, int ox, int oy, bool edge_fill,
                               float fill) {
  int nx, ny;
  std::vector<float> d = dense_unrotated(padded, W, H, rotated, &nx, &ny);
  if (edge_fill) {
    double s = 0; long n = 0;
    for (int x = 0; x < nx; ++x) { s += d[x]; s += d[(size_t)(ny - 1) * nx + x]; n += 2; }
    for (int y = 1; y < ny - 1; ++y) { s += d[(size_t)y * nx]; s += d[(size_t)y * nx + nx - 1]; n += 2; }
    fill = (float)(s / (double)(float)n);
  }
  std::vector<float> o((size_t)ox * oy);
  for (int jj = 0; jj < oy; ++jj) {
    const int j = ny / 2 + (jj - oy / 2);
    for (int ii = 0; ii < ox; ++ii) {
      const int i = nx / 2 + (ii - ox / 2);
      o[(size_t)jj * ox + ii] = (i < 0 || i >= nx || j < 0 || j >= ny) ? fill : d[(size_t)j * nx + i];
    }
  }
  return o;
}

// The job arguments of match_template, in the reference's order, names and types
// (match_template.cpp:346-387; type string "ttffffffffffifffffbfftttttttttftiiiitttfbi", :237-278),
// with the defaults of its interactive input (:158-222).
struct MatchTemplateParams {
  std::string input_search_images = "image_stack.mrc";
  std::string input_reconstruction = "reconstruction.mrc";
  float pixel_size = 1.0f;
  float voltage_kV = 300.0f;
  float spherical_aberration_mm = 2.7f;
  float amplitude_contrast = 0.07f;
  float defocus1 = 10000.0f;
  float defocus2 = 10000.0f;
  float defocus_angle = 0.0f;
  float low_resolution_limit = 300.0f;
  float high_resolution_limit_search = 8.0f;
  float angular_step = 0.0f;
  int best_parameters_to_keep = 20;
  float defocus_search_range = 500.0f;
  float defocus_step = 50.0f;
  float pixel_size_search_range = 0.1f;
  float pixel_size_step = 0.01f;
  float padding = 1.0f;
  bool ctf_refinement = false;
  float particle_radius_angstroms = 0.0f;
  float phase_shift = 0.0f;
  std::string mip_output_file = "mip.mrc";
  std::string best_psi_output_file = "psi.mrc";
  std::string best_theta_output_file = "theta.mrc";
  std::string best_phi_output_file = "phi.mrc";
  std::string best_defocus_output_file = "defocus.mrc";
  std::string best_pixel_size_output_file = "pixel_size.mrc";
  std::string scaled_mip_output_file = "mip_scaled.mrc";
  std::string correlation_avg_output_file = "corr_average.mrc";
  std::string my_symmetry = "C1";
  float in_plane_angular_step = 0.0f;
  std::string output_histogram_file = "histogram.txt";
  int first_search_position = -1;
  int last_search_position = -1;
  int image_number_for_gui = 0;
  int number_of_jobs_per_image_in_gui = 0;
  std::string correlation_std_output_file = "corr_variance.mrc";
  std::string directory_for_results = "/dev/null";
  std::string result_output_filename = "/dev/null";
  float min_peak_radius = 10.0f;
  bool use_gpu = false;
  int max_threads = 1;
};

constexpr const char* kJobTypes = "ttffffffffffifffffbfftttttttttftiiiitttfbi";

bool parse_bool(const std::string& v) {
  std::string l;
  for (char c : v) l += (char)std::tolower((unsigned char)c);
  if (l == "1" || l == "y" || l == "yes" || l == "true") return true;
  if (l == "0" || l == "n" || l == "no" || l == "false") return false;
  throw std::runtime_error("not a yes/no value: '" + v + "'");
}

float parse_float(const std::string& v) {
  size_t used = 0;
  const float x = std::stof(v, &used);
  if (used != v.size()) throw std::runtime_error("not a number: '" + v + "'");
  return x;
}

int parse_int(const std::string& v) {
  size_t used = 0;
  const int x = std::stoi(v, &used);
  if (used != v.size()) throw std::runtime_error("not an integer: '" + v + "'");
  return x;
}

// The 42 job arguments (RunJob order) -> parameters.
MatchTemplateParams from_job_arguments(const std::vector<std::string>& v) {
  if (v.size() != std::strlen(kJobTypes))
    throw std::runtime_error("--job-args takes " + std::to_string(std::strlen(kJobTypes)) + " values, got " +
                             std::to_string(v.size()));
  MatchTemplateParams p;
  void* slot[] = {&p.input_search_images, &p.input_reconstruction, &p.pixel_size, &p.voltage_kV,
                  &p.spherical_aberration_mm, &p.amplitude_contrast, &p.defocus1, &p.defocus2, &p.defocus_angle,
                  &p.low_resolution_limit, &p.high_resolution_limit_search, &p.angular_step,
                  &p.best_parameters_to_keep, &p.defocus_search_range, &p.defocus_step, &p.pixel_size_search_range,
                  &p.pixel_size_step, &p.padding, &p.ctf_refinement, &p.particle_radius_angstroms, &p.phase_shift,
                  &p.mip_output_file, &p.best_psi_output_file, &p.best_theta_output_file, &p.best_phi_output_file,
                  &p.best_defocus_output_file, &p.best_pixel_size_output_file, &p.scaled_mip_output_file,
                  &p.correlation_avg_output_file, &p.my_symmetry, &p.in_plane_angular_step, &p.output_histogram_file,
                  &p.first_search_position, &p.last_search_position, &p.image_number_for_gui,
                  &p.number_of_jobs_per_image_in_gui, &p.correlation_std_output_file, &p.directory_for_results,
                  &p.result_output_filename, &p.min_peak_radius, &p.use_gpu, &p.max_threads};
  static_assert(sizeof(slot) / sizeof(slot[0]) == 42, "one slot per job argument");
  for (size_t k = 0; k < v.size(); ++k) {
    switch (kJobTypes[k]) {
      case 't': *static_cast<std::string*>(slot[k]) = v[k]; break;
      case 'f': *static_cast<float*>(slot[k]) = parse_float(v[k]); break;
      case 'i': *static_cast<int*>(slot[k]) = parse_int(v[k]); break;
      case 'b': *static_cast<bool*>(slot[k]) = parse_bool(v[k]); break;
    }
  }
  return p;
}

// The interactive questions of MatchTemplateApp::DoInteractiveUserInput (match_template.cpp:184-214,
// with the two GPU questions of a GPU build), one answer per line of `in`; an empty line takes the
// default. Questions go to stderr.
MatchTemplateParams from_interactive(std::istream& in) {
  MatchTemplateParams p;
  struct Q {
    const char* text;
    char type;
    void* slot;
    const char* def;
    double lo, hi;  // accepted range of a number (lo > hi: none)
  };
  const Q qs[] = {
      {"Input images to be searched", 't', &p.input_search_images, "image_stack.mrc", 1, 0},
      {"Input template reconstruction", 't', &p.input_reconstruction, "reconstruction.mrc", 1, 0},
      {"Output MIP file", 't', &p.mip_output_file, "mip.mrc", 1, 0},
      {"Output Scaled MIP file", 't', &p.scaled_mip_output_file, "mip_scaled.mrc", 1, 0},
      {"Output psi file", 't', &p.best_psi_output_file, "psi.mrc", 1, 0},
      {"Output theta file", 't', &p.best_theta_output_file, "theta.mrc", 1, 0},
      {"Output phi file", 't', &p.best_phi_output_file, "phi.mrc", 1, 0},
      {"Output defocus file", 't', &p.best_defocus_output_file, "defocus.mrc", 1, 0},
      {"Output pixel size file", 't', &p.best_pixel_size_output_file, "pixel_size.mrc", 1, 0},
      {"Correlation average value", 't', &p.correlation_avg_output_file, "corr_average.mrc", 1, 0},
      {"Correlation variance output file", 't', &p.correlation_std_output_file, "corr_variance.mrc", 1, 0},
      {"Output histogram of correlation values", 't', &p.output_histogram_file, "histogram.txt", 1, 0},
      {"Pixel size of images (A)", 'f', &p.pixel_size, "1.0", 0.0, 1e30},
      {"Beam energy (keV)", 'f', &p.voltage_kV, "300.0", 0.0, 1e30},
      {"Spherical aberration (mm)", 'f', &p.spherical_aberration_mm, "2.7", 1, 0},
      {"Amplitude contrast", 'f', &p.amplitude_contrast, "0.07", 0.0, 1.0},
      {"Defocus1 (angstroms)", 'f', &p.defocus1, "10000", 0.0, 1e30},
      {"Defocus2 (angstroms)", 'f', &p.defocus2, "10000", 0.0, 1e30},
      {"Defocus Angle (degrees)", 'f', &p.defocus_angle, "0.0", 1, 0},
      {"Phase Shift (degrees)", 'f', &p.phase_shift, "0.0", 1, 0},
      {"High resolution limit (A)", 'f', &p.high_resolution_limit_search, "8.0", 0.0, 1e30},
      {"Out of plane angular step (0.0 = set automatically)", 'f', &p.angular_step, "0.0", 0.0, 1e30},
      {"In plane angular step (0.0 = set automatically)", 'f', &p.in_plane_angular_step, "0.0", 0.0, 1e30},
      {"Defocus search range (A)", 'f', &p.defocus_search_range, "500.0", 0.0, 1e30},
      {"Defocus step (A) (0.0 = no search)", 'f', &p.defocus_step, "50.0", 0.0, 1e30},
      {"Pixel size search range (A)", 'f', &p.pixel_size_search_range, "0.1", 0.0, 1e30},
      {"Pixel size step (A) (0.0 = no search)", 'f', &p.pixel_size_step, "0.01", 0.0, 1e30},
      {"Padding factor", 'f', &p.padding, "1.0", 1.0, 2.0},
      {"Mask radius for global search (A) (0.0 = max)", 'f', &p.particle_radius_angstroms, "0.0", 0.0, 1e30},
      {"Template symmetry", 't', &p.my_symmetry, "C1", 1, 0},
      {"Use GPU", 'b', &p.use_gpu, "No", 1, 0},
      {"Max. threads to use for calculation", 'i', &p.max_threads, "1", 1.0, 1e30},
  };
  for (const Q& q : qs) {
    std::fprintf(stderr, "%s [%s] : ", q.text, q.def);
    std::string line;
    if (!std::getline(in, line)) line.clear();
    while (!line.empty() && (line.back() == '\r' || line.back() == ' ')) line.pop_back();
    const std::string v = line.empty() ? std::string(q.def) : line;
    double num = 0.0;
    switch (q.type) {
      case 't': *static_cast<std::string*>(q.slot) = v; break;
      case 'f': num = *static_cast<float*>(q.slot) = parse_float(v); break;
      case 'i': num = *static_cast<int*>(q.slot) = parse_int(v); break;
      case 'b': *static_cast<bool*>(q.slot) = parse_bool(v); break;
    }
    if ((q.type == 'f' || q.type == 'i') && q.lo <= q.hi && (num < q.lo || num > q.hi))
      throw std::runtime_error(std::string(q.text) + ": " + v + " is outside [" + std::to_string(q.lo) + ", " +
                               std::to_string(q.hi) + "]");
    std::fprintf(stderr, "%s\n", v.c_str());
  }
  return p;
}

// Options of this program that the reference's job arguments do not carry.
struct DriverOptions {
  int device = 0, batch = 0, graph = 0;
  bool dry_run = false;
  std::string worker_record;  // empty: none
  std::string peaks_file;     // empty: none
};

void usage() {
  std::fprintf(stderr,
               "usage: match_template_hip --image F --template F --out PREFIX [--pixel-size 1] [--voltage 300]\n"
               "  [--cs 2.7] [--amplitude-contrast 0.07] [--defocus1 A] [--defocus2 A] [--defocus-angle 0]\n"
               "  [--phase-shift 0] [--angular-step 10] [--in-plane-step 10] [--defocus-range 0] [--defocus-step 0]\n"
               "  [--pixel-size-range 0] [--pixel-size-step 0] [--padding 1] [--symmetry C1] [--min-peak-radius 10]\n"
               "  [--high-resolution-limit 8] [--first-search-position -1] [--last-search-position -1]\n"
               "  [--device 0] [--batch 0] [--graph 0|1] [--particle-radius 0] [--worker-record FILE] [--dry-run 1]\n"
               "or:    match_template_hip [--device N] [--batch N] [--dry-run] --job-args <the 42 job arguments of\n"
               "       match_template, in its order>\n"
               "or:    match_template_hip [--device N] [--batch N] [--dry-run] --interactive   (answers on stdin)\n");
}

// The local run of MatchTemplateApp::DoCalculation on one GPU.
int run(MatchTemplateParams prm, const DriverOptions& opt) {
  const float pixel_size = prm.pixel_size, kv = prm.voltage_kV, cs = prm.spherical_aberration_mm;
  const float amp = prm.amplitude_contrast, df1 = prm.defocus1, df2 = prm.defocus2;
  const float dfa = prm.defocus_angle, phase = prm.phase_shift;
  // match_template.cpp:657-676: a mask radius of 200 A unless one is given; the steps from it
  const float mask_radius_search = prm.particle_radius_angstroms < 1.0f ? 200.0f : prm.particle_radius_angstroms;
  float ang_step = prm.angular_step;
  if (ang_step <= 0) ang_step = (float)(360.0 * prm.high_resolution_limit_search / 3.14159265359 / mask_radius_search);
  float psi_step;
  if (prm.in_plane_angular_step <= 0) {
    psi_step = pixel_size / mask_radius_search / (3.14159265359f / 180.);
    psi_step = 360.0 / int(360.0 / psi_step + 0.5);
  } else {
    psi_step = prm.in_plane_angular_step;
  }
  float def_range = prm.defocus_search_range, def_step = prm.defocus_step;
  float pix_range = prm.pixel_size_search_range, pix_step = prm.pixel_size_step;
  const float padding = prm.padding, min_peak_radius = prm.min_peak_radius;
  const int device = opt.device, batch = opt.batch, graph = opt.graph;
  const std::string sym = prm.my_symmetry;
  const int sym_n = sym.size() > 1 ? std::stoi(sym.substr(1)) : 0;
  const int npos = tm_search_grid(sym[0], sym_n, ang_step, nullptr, 0);
  if (npos <= 0) throw std::runtime_error("no search positions for symmetry " + sym);
  // a local run searches every position (match_template.cpp:735-739); a worker's range applies
  // when both ends are given
  const int first = prm.first_search_position >= 0 && prm.last_search_position >= 0 ? prm.first_search_position : 0;
  const int last = prm.first_search_position >= 0 && prm.last_search_position >= 0
                       ? std::min(prm.last_search_position, npos - 1)
                       : npos - 1;
  if (first > last) throw std::runtime_error("empty search position range");
  const int n_psi = tm_psi_list(0.0f, 360.0f, psi_step, nullptr, 0);
  if (def_step <= 0.0) { def_range = 0.0f; def_step = 100.0f; }
  const float r = def_range / def_step;
  const int ndh = r > 0 ? int(r + 0.5) : int(r - 0.5);
  // match_template.cpp:755-762
  if (pix_step <= 0.0f) { pix_range = 0.0f; pix_step = 0.02f; }
  const float rs = pix_range / pix_step;
  const int nsh = rs > 0 ? int(rs + 0.5) : int(rs - 0.5);
  if (opt.dry_run) {
    std::printf("angular_step=%.9g\npsi_step=%.9g\nsearch_positions=%d\nfirst_search_position=%d\n"
                "last_search_position=%d\npsi_per_position=%d\ndefocus_planes=%d\npixel_size_planes=%d\n"
                "defocus_step=%.9g\npixel_size_step=%.9g\nmask_radius_search=%.9g\n",
                ang_step, psi_step, npos, first, last, n_psi, 2 * ndh + 1, 2 * nsh + 1, def_step, pix_step,
                mask_radius_search);
    std::printf("input_search_images=%s\ninput_reconstruction=%s\nmip_output_file=%s\nscaled_mip_output_file=%s\n"
                "best_psi_output_file=%s\nbest_theta_output_file=%s\nbest_phi_output_file=%s\n"
                "best_defocus_output_file=%s\nbest_pixel_size_output_file=%s\ncorrelation_avg_output_file=%s\n"
                "correlation_std_output_file=%s\noutput_histogram_file=%s\nsymmetry=%s\npixel_size=%.9g\n"
                "voltage_kV=%.9g\nspherical_aberration_mm=%.9g\namplitude_contrast=%.9g\ndefocus1=%.9g\n"
                "defocus2=%.9g\ndefocus_angle=%.9g\nphase_shift=%.9g\npadding=%.9g\nmin_peak_radius=%.9g\n"
                "particle_radius_angstroms=%.9g\nuse_gpu=%d\nmax_threads=%d\n",
                prm.input_search_images.c_str(), prm.input_reconstruction.c_str(), prm.mip_output_file.c_str(),
                prm.scaled_mip_output_file.c_str(), prm.best_psi_output_file.c_str(), prm.best_theta_output_file.c_str(),
                prm.best_phi_output_file.c_str(), prm.best_defocus_output_file.c_str(),
                prm.best_pixel_size_output_file.c_str(), prm.correlation_avg_output_file.c_str(),
                prm.correlation_std_output_file.c_str(), prm.output_histogram_file.c_str(), sym.c_str(), pixel_size,
                kv, cs, amp, df1, df2, dfa, phase, padding, min_peak_radius, prm.particle_radius_angstroms,
                (int)prm.use_gpu, prm.max_threads);
    return 0;
  }
  if (!prm.use_gpu)
    std::fprintf(stderr, "note: this program always runs the search on the GPU (Use GPU = No is ignored)\n");
  {
    Vol img = read_mrc(prm.input_search_images);
    Vol vol = read_mrc(prm.input_reconstruction);
    if (img.nz != 1 || vol.nx != vol.ny || vol.nx != vol.nz) throw std::runtime_error("need a 2-D image and a cubic template");
    const int T = vol.nx;
    int W, H, rot, ncurve, maxpad;
    if (tm_prepare_sizes(img.nx, img.ny, T, &W, &H, &rot, &ncurve, &maxpad)) throw std::runtime_error(tm_prepare_last_error());
    std::vector<float> spec((size_t)H * (W / 2 + 1) * 2), cx(ncurve), cy(ncurve);
    if (tm_prepare_input_image(device, img.v.data(), img.nx, img.ny, T, spec.data(), cx.data(), cy.data()))
      throw std::runtime_error(tm_prepare_last_error());
    int P = 0;
    if (tm_prepare_template(device, vol.v.data(), T, padding, nullptr, &P)) throw std::runtime_error(tm_prepare_last_error());
    std::vector<float> tpl((size_t)P * P * (P / 2 + 1) * 2);
    if (tm_prepare_template(device, vol.v.data(), T, padding, tpl.data(), &P)) throw std::runtime_error(tm_prepare_last_error());
    std::vector<float> grid(2 * (size_t)npos);
    tm_search_grid(sym[0], sym_n, ang_step, grid.data(), npos);
    const double sqrt_px = sqrt((double)(W * H));
    const float hstep = (22.5f - -12.5f) / float(512);
    cistem_amd::TemplateMatchingCore core(device);
    core.Init(tpl.data(), P, spec.data(), W, H, T, pix_step, def_step, 360.0f, 0.0f, psi_step, grid.data(), npos,
              -12.5 / sqrt_px, hstep / sqrt_px, 512, maxpad, first, last, batch);
    if (graph) core.SetGraph(true);  // each plane replayed from one captured hipGraph
    std::vector<float> pf((size_t)T * (T / 2 + 1) * 2);
    // pixel-size planes, then defocus planes (match_template.cpp:837-914)
    for (int si = -nsh; si <= nsh; ++si) {
      if (nsh) {
        const float factor = (pixel_size + float(si) * pix_step) / pixel_size;
        if (tm_prepare_template_scaled(device, vol.v.data(), T, padding, factor, tpl.data(), &P))
          throw std::runtime_error(tm_prepare_last_error());
        core.SetTemplate(tpl.data());
      }
      for (int di = -ndh; di <= ndh; ++di) {
        if (tm_projection_filter(T, cx.data(), cy.data(), ncurve, kv, cs, amp, df1 + float(di) * def_step,
                                 df2 + float(di) * def_step, dfa, pixel_size, phase, pf.data()))
          throw std::runtime_error(tm_prepare_last_error());
        const long long n = core.RunInnerLoop(pf.data(), si, di);
        std::fprintf(stderr, "pixel-size plane %d, defocus plane %d: %lld CCFs\n", si, di, n);
      }
    }
    // the count the scaling divides by: positions x psi x defocus planes, without the pixel-size
    // planes, as the reference's total_correlation_positions (match_template.cpp:730-764)
    const long long total = core.total_number_of_cccs_calculated / (2 * nsh + 1);
    const size_t np = (size_t)(W + 2) * H;
    std::vector<float> mip(np), psi(np), theta(np), phi(np), defo(np), pix(np);
    std::vector<long> hist(512, 0);
    core.CopyResults(mip.data(), psi.data(), theta.data(), phi.data(), defo.data(), pix.data(), nullptr, nullptr,
                     hist.data());
    std::vector<float> mips(np), z(np), avg(np), sd(np);
    core.Finalize(total, mips.data(), z.data(), avg.data(), sd.data());
    const int ox = img.nx, oy = img.ny;
    const bool r90 = rot != 0;
    if (!opt.worker_record.empty()) {
      // the record a worker sends its master (match_template.cpp:1371-1531): the 8 maps in the
      // micrograph's orientation, remove_npix_from_edge = myroundint(particle radius / pixel size)
      // (match_template.cpp:494), the CCFs this run computed, sqrt(W H)
      std::vector<double> s1(np), s2(np);
      std::vector<long> hist_unused(512, 0);
      core.CopyResults(mip.data(), psi.data(), theta.data(), phi.data(), defo.data(), pix.data(), s1.data(), s2.data(),
                       hist_unused.data());
      int nx = 0, ny = 0;
      std::vector<std::vector<float>> maps;
      maps.push_back(dense_unrotated(mip.data(), W, H, r90, &nx, &ny));
      for (const std::vector<float>* m : {&psi, &theta, &phi, &defo, &pix})
        maps.push_back(dense_unrotated(m->data(), W, H, r90, &nx, &ny));
      maps.push_back(dense_unrotated(s1.data(), W, H, r90, &nx, &ny));
      maps.push_back(dense_unrotated(s2.data(), W, H, r90, &nx, &ny));
      const float* ptrs[TM_RESULT_MAPS];
      for (int k = 0; k < TM_RESULT_MAPS; ++k) ptrs[k] = maps[k].data();
      const float q = prm.particle_radius_angstroms / pixel_size;
      const int remove = q > 0 ? (int)(q + 0.5f) : (int)(q - 0.5f);
      std::vector<long long> h64(hist.begin(), hist.end());
      const long long nf = tm_result_size(ox, oy, 512);
      std::vector<float> rec((size_t)nf);
      if (tm_result_pack(ptrs, nx, ny, ox, oy, remove, h64.data(), 512, (float)core.total_number_of_cccs_calculated,
                         std::sqrt((double)W * (double)H), rec.data(), nf) != TM_OK)
        throw std::runtime_error(tm_result_last_error());
      std::ofstream wf(opt.worker_record, std::ios::binary);
      wf.write(reinterpret_cast<const char*>(rec.data()), (std::streamsize)(rec.size() * sizeof(float)));
      std::fprintf(stderr, "worker record: %lld floats -> %s\n", nf, opt.worker_record.c_str());
    }
    auto mip_o = to_original(mips.data(), W, H, r90, ox, oy, true, 0.f);
    auto z_o = to_original(z.data(), W, H, r90, ox, oy, true, 0.f);
    auto avg_o = to_original(avg.data(), W, H, r90, ox, oy, true, 0.f);
    auto sd_o = to_original(sd.data(), W, H, r90, ox, oy, true, 0.f);
    auto psi_o = to_original(psi.data(), W, H, r90, ox, oy, false, 0.f);
    auto theta_o = to_original(theta.data(), W, H, r90, ox, oy, false, 0.f);
    auto phi_o = to_original(phi.data(), W, H, r90, ox, oy, false, 0.f);
    auto def_o = to_original(defo.data(), W, H, r90, ox, oy, false, 0.f);
    auto pix_o = to_original(pix.data(), W, H, r90, ox, oy, false, 0.f);
    write_mrc(prm.mip_output_file, mip_o.data(), ox, oy, pixel_size);
    write_mrc(prm.scaled_mip_output_file, z_o.data(), ox, oy, pixel_size);
    write_mrc(prm.correlation_avg_output_file, avg_o.data(), ox, oy, pixel_size);
    write_mrc(prm.correlation_std_output_file, sd_o.data(), ox, oy, pixel_size);
    write_mrc(prm.best_psi_output_file, psi_o.data(), ox, oy, pixel_size);
    write_mrc(prm.best_theta_output_file, theta_o.data(), ox, oy, pixel_size);
    write_mrc(prm.best_phi_output_file, phi_o.data(), ox, oy, pixel_size);
    write_mrc(prm.best_defocus_output_file, def_o.data(), ox, oy, pixel_size);
    write_mrc(prm.best_pixel_size_output_file, pix_o.data(), ox, oy, pixel_size);

    // expected threshold, master form (match_template.cpp:1799)
    const float thr = (float)tm_expected_threshold(ox, oy, total);
    // histogram file (match_template.cpp:1333-1364), bins 0..511
    {
      std::vector<long long> h64(hist.begin(), hist.end());
      std::vector<float> snr(512);
      std::vector<double> surv(512, 0.0), expect(512, 0.0);
      if (tm_histogram_table(h64.data(), 512, -12.5f, hstep, sqrt_px, total, snr.data(), surv.data(), expect.data()) !=
          TM_OK)
        throw std::runtime_error(std::string("tm_histogram_table: ") + tm_result_last_error());
      FILE* hf = std::fopen(prm.output_histogram_file.c_str(), "w");
      if (!hf) throw std::runtime_error("cannot write " + prm.output_histogram_file);
      std::fprintf(hf, "# Expected threshold = %.2f\n", thr);
      std::fprintf(hf, "# SNR, histogram, survival histogram, random survival histogram\n");
      for (int i = 0; i < 512; ++i) std::fprintf(hf, "%f %ld %f %f\n", snr[i], hist[i], surv[i], expect[i]);
      std::fclose(hf);
    }
    // peaks (match_template.cpp:1839-1980): integer maximum of the z-score map away from a
    // T/4 + 1 border; mask a disc of min_peak_radius; stop below the threshold or at 1000 peaks
    if (opt.peaks_file.empty()) {
      // the reference's local run writes no peak list
    } else if (!r90 && W == ox && H == oy) {
      // the search geometry is the image's: the peak search runs on the device
      std::vector<tm_peak_info> pk(1000);
      const int found = core.FindPeaks(total, thr, T / 4 + 1, min_peak_radius, pixel_size, pk.data(), 1000);
      FILE* pf2 = std::fopen(opt.peaks_file.c_str(), "w");
      std::fprintf(pf2, "# x_A y_A psi theta phi defocus pixel_size peak_height\n");
      for (int q = 0; q < found; ++q) {
        const tm_peak_info& r = pk[q];
        std::fprintf(pf2, "%f %f %f %f %f %f %f %f\n", r.x_pos, r.y_pos, r.psi, r.theta, r.phi, r.defocus,
                     r.pixel_size, r.peak_height);
      }
      std::fclose(pf2);
      std::fprintf(stderr, "%d peaks above the expected threshold %.3f\n", found, thr);
    } else {
      std::vector<float> zz = z_o;
      const int edge = T / 4 + 1;
      const int rad = (int)min_peak_radius;
      const float rad2 = powf(float(rad), 2);
      FILE* pf2 = std::fopen(opt.peaks_file.c_str(), "w");
      std::fprintf(pf2, "# x_A y_A psi theta phi defocus pixel_size peak_height\n");
      int found = 0;
      while (found < 1000) {
        float best = -FLT_MAX;
        int bx = 0, by = 0;
        for (int y = edge; y < oy - edge; ++y)
          for (int x = edge; x < ox - edge; ++x)
            if (zz[(size_t)y * ox + x] > best) { best = zz[(size_t)y * ox + x]; bx = x; by = y; }
        if (best < thr) break;
        const size_t at = (size_t)by * ox + bx;
        std::fprintf(pf2, "%f %f %f %f %f %f %f %f\n", bx * pixel_size, by * pixel_size, psi_o[at], theta_o[at],
                     phi_o[at], def_o[at], pix_o[at], z_o[at]);
        for (int j = std::max(by - rad, 0); j < std::min(by + rad, oy); ++j)
          for (int i = std::max(bx - rad, 0); i < std::min(bx + rad, ox); ++i) {
            const float dx = float(i) - bx, dy = float(j) - by;
            if (dx * dx + dy * dy <= rad2) zz[(size_t)j * ox + i] = -FLT_MAX;
          }
        ++found;
      }
      std::fclose(pf2);
      std::fprintf(stderr, "%d peaks above the expected threshold %.3f\n", found, thr);
    }
    std::fprintf(stderr, "total CCFs %lld\n", total);
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    MatchTemplateParams prm;
    DriverOptions opt;
    std::vector<std::string> args(argv + 1, argv + argc);
    // options of this program, in any mode
    auto take = [&](const char* name, std::string* value) {
      for (size_t i = 0; i < args.size(); ++i)
        if (args[i] == name) {
          if (value) {
            if (i + 1 >= args.size()) throw std::runtime_error(std::string(name) + " needs a value");
            *value = args[i + 1];
            args.erase(args.begin() + i, args.begin() + i + 2);
          } else {
            args.erase(args.begin() + i);
          }
          return true;
        }
      return false;
    };
    std::string v;
    if (!args.empty() && args[0] == "--job-args") {
      prm = from_job_arguments(std::vector<std::string>(args.begin() + 1, args.end()));
      args.clear();
    } else {
      size_t ja = std::find(args.begin(), args.end(), std::string("--job-args")) - args.begin();
      if (ja < args.size()) {
        std::vector<std::string> job(args.begin() + ja + 1, args.end());
        args.erase(args.begin() + ja, args.end());
        prm = from_job_arguments(job);
        if (take("--device", &v)) opt.device = parse_int(v);
        if (take("--batch", &v)) opt.batch = parse_int(v);
        opt.dry_run = take("--dry-run", nullptr);
        if (!args.empty()) throw std::runtime_error("unknown argument before --job-args: " + args[0]);
      } else if (std::find(args.begin(), args.end(), std::string("--interactive")) != args.end()) {
        take("--interactive", nullptr);
        if (take("--device", &v)) opt.device = parse_int(v);
        if (take("--batch", &v)) opt.batch = parse_int(v);
        opt.dry_run = take("--dry-run", nullptr);
        if (!args.empty()) throw std::runtime_error("unknown argument: " + args[0]);
        prm = from_interactive(std::cin);
      } else {
        // named options (this program's own form): the search of a prefix's outputs
        std::map<std::string, std::string> a;
        for (size_t i = 0; i + 1 < args.size(); i += 2) {
          if (args[i].compare(0, 2, "--") != 0) { usage(); return 2; }
          a[args[i].substr(2)] = args[i + 1];
        }
        if (args.size() % 2 || !a.count("image") || !a.count("template") || !a.count("out")) { usage(); return 2; }
        auto f = [&](const char* k, double d) { return a.count(k) ? std::stod(a[k]) : d; };
        const std::string out = a["out"];
        prm.input_search_images = a["image"];
        prm.input_reconstruction = a["template"];
        prm.pixel_size = f("pixel-size", 1.0);
        prm.voltage_kV = f("voltage", 300.0);
        prm.spherical_aberration_mm = f("cs", 2.7);
        prm.amplitude_contrast = f("amplitude-contrast", 0.07);
        prm.defocus1 = f("defocus1", 15000.0);
        prm.defocus2 = f("defocus2", prm.defocus1);
        prm.defocus_angle = f("defocus-angle", 0.0);
        prm.phase_shift = f("phase-shift", 0.0);
        prm.high_resolution_limit_search = f("high-resolution-limit", 8.0);
        prm.angular_step = f("angular-step", 10.0);
        prm.in_plane_angular_step = f("in-plane-step", 10.0);
        prm.defocus_search_range = f("defocus-range", 0.0);
        prm.defocus_step = f("defocus-step", 0.0);
        prm.pixel_size_search_range = f("pixel-size-range", 0.0);
        prm.pixel_size_step = f("pixel-size-step", 0.0);
        prm.padding = f("padding", 1.0);
        prm.min_peak_radius = f("min-peak-radius", 10.0);
        prm.particle_radius_angstroms = f("particle-radius", 0.0);
        prm.first_search_position = (int)f("first-search-position", -1);
        prm.last_search_position = (int)f("last-search-position", -1);
        prm.my_symmetry = a.count("symmetry") ? a["symmetry"] : "C1";
        prm.use_gpu = true;
        prm.mip_output_file = out + "_mip.mrc";
        prm.scaled_mip_output_file = out + "_scaled_mip.mrc";
        prm.correlation_avg_output_file = out + "_avg.mrc";
        prm.correlation_std_output_file = out + "_std.mrc";
        prm.best_psi_output_file = out + "_psi.mrc";
        prm.best_theta_output_file = out + "_theta.mrc";
        prm.best_phi_output_file = out + "_phi.mrc";
        prm.best_defocus_output_file = out + "_defocus.mrc";
        prm.best_pixel_size_output_file = out + "_pixel_size.mrc";
        prm.output_histogram_file = out + "_histogram.txt";
        opt.device = (int)f("device", 0);
        opt.batch = (int)f("batch", 0);
        opt.graph = (int)f("graph", 0);
        opt.dry_run = f("dry-run", 0) != 0;
        opt.peaks_file = out + "_peaks.txt";
        if (a.count("worker-record")) opt.worker_record = a["worker-record"];
      }
    }
    return run(prm, opt);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "match_template_hip: %s\n", e.what());
    return 1;
  }
}
