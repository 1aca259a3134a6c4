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
#include <fstream>
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

void usage() {
  std::fprintf(stderr,
               "usage: match_template_hip --image F --template F --out PREFIX [--pixel-size 1] [--voltage 300]\n"
               "  [--cs 2.7] [--amplitude-contrast 0.07] [--defocus1 A] [--defocus2 A] [--defocus-angle 0]\n"
               "  [--phase-shift 0] [--angular-step 10] [--in-plane-step 10] [--defocus-range 0] [--defocus-step 0]\n"
               "  [--pixel-size-range 0] [--pixel-size-step 0] [--padding 1] [--symmetry C1] [--min-peak-radius 10]\n"
               "  [--device 0] [--batch 0] [--graph 0|1] [--particle-radius 0] [--worker-record FILE]\n");
}

}  // namespace

int main(int argc, char** argv) {
  std::map<std::string, std::string> a;
  for (int i = 1; i + 1 < argc; i += 2) {
    if (std::strncmp(argv[i], "--", 2) != 0) { usage(); return 2; }
    a[argv[i] + 2] = argv[i + 1];
  }
  if (!a.count("image") || !a.count("template") || !a.count("out")) { usage(); return 2; }
  auto f = [&](const char* k, double d) { return a.count(k) ? std::stod(a[k]) : d; };
  try {
    const float pixel_size = f("pixel-size", 1.0), kv = f("voltage", 300.0), cs = f("cs", 2.7);
    const float amp = f("amplitude-contrast", 0.07), df1 = f("defocus1", 15000.0), df2 = f("defocus2", df1);
    const float dfa = f("defocus-angle", 0.0), phase = f("phase-shift", 0.0);
    const float ang_step = f("angular-step", 10.0), psi_step = f("in-plane-step", 10.0);
    float def_range = f("defocus-range", 0.0), def_step = f("defocus-step", 0.0);
    float pix_range = f("pixel-size-range", 0.0), pix_step = f("pixel-size-step", 0.0);
    const float padding = f("padding", 1.0), min_peak_radius = f("min-peak-radius", 10.0);
    const int device = (int)f("device", 0), batch = (int)f("batch", 0), graph = (int)f("graph", 0);
    const std::string sym = a.count("symmetry") ? a["symmetry"] : "C1";
    const std::string out = a["out"];

    Vol img = read_mrc(a["image"]);
    Vol vol = read_mrc(a["template"]);
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
    const int npos = tm_search_grid(sym[0], sym.size() > 1 ? std::stoi(sym.substr(1)) : 0, ang_step, nullptr, 0);
    std::vector<float> grid(2 * (size_t)npos);
    tm_search_grid(sym[0], sym.size() > 1 ? std::stoi(sym.substr(1)) : 0, ang_step, grid.data(), npos);

    if (def_step <= 0.0) { def_range = 0.0f; def_step = 100.0f; }
    const float r = def_range / def_step;
    const int ndh = r > 0 ? int(r + 0.5) : int(r - 0.5);
    // match_template.cpp:755-762
    if (pix_step <= 0.0f) { pix_range = 0.0f; pix_step = 0.02f; }
    const float rs = pix_range / pix_step;
    const int nsh = rs > 0 ? int(rs + 0.5) : int(rs - 0.5);
    const double sqrt_px = sqrt((double)(W * H));
    const float hstep = (22.5f - -12.5f) / float(512);
    cistem_amd::TemplateMatchingCore core(device);
    core.Init(tpl.data(), P, spec.data(), W, H, T, pix_step, def_step, 360.0f, 0.0f, psi_step, grid.data(), npos,
              -12.5 / sqrt_px, hstep / sqrt_px, 512, maxpad, 0, npos - 1, batch);
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
    if (a.count("worker-record")) {
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
      const float q = (float)f("particle-radius", 0.0) / pixel_size;
      const int remove = q > 0 ? (int)(q + 0.5f) : (int)(q - 0.5f);
      std::vector<long long> h64(hist.begin(), hist.end());
      const long long nf = tm_result_size(ox, oy, 512);
      std::vector<float> rec((size_t)nf);
      if (tm_result_pack(ptrs, nx, ny, ox, oy, remove, h64.data(), 512, (float)core.total_number_of_cccs_calculated,
                         std::sqrt((double)W * (double)H), rec.data(), nf) != TM_OK)
        throw std::runtime_error(tm_result_last_error());
      std::ofstream wf(a["worker-record"], std::ios::binary);
      wf.write(reinterpret_cast<const char*>(rec.data()), (std::streamsize)(rec.size() * sizeof(float)));
      std::fprintf(stderr, "worker record: %lld floats -> %s\n", nf, a["worker-record"].c_str());
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
    write_mrc(out + "_mip.mrc", mip_o.data(), ox, oy, pixel_size);
    write_mrc(out + "_scaled_mip.mrc", z_o.data(), ox, oy, pixel_size);
    write_mrc(out + "_avg.mrc", avg_o.data(), ox, oy, pixel_size);
    write_mrc(out + "_std.mrc", sd_o.data(), ox, oy, pixel_size);
    write_mrc(out + "_psi.mrc", psi_o.data(), ox, oy, pixel_size);
    write_mrc(out + "_theta.mrc", theta_o.data(), ox, oy, pixel_size);
    write_mrc(out + "_phi.mrc", phi_o.data(), ox, oy, pixel_size);
    write_mrc(out + "_defocus.mrc", def_o.data(), ox, oy, pixel_size);
    write_mrc(out + "_pixel_size.mrc", pix_o.data(), ox, oy, pixel_size);

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
      FILE* hf = std::fopen((out + "_histogram.txt").c_str(), "w");
      std::fprintf(hf, "# Expected threshold = %.2f\n", thr);
      std::fprintf(hf, "# SNR, histogram, survival histogram, random survival histogram\n");
      for (int i = 0; i < 512; ++i) std::fprintf(hf, "%f %ld %f %f\n", snr[i], hist[i], surv[i], expect[i]);
      std::fclose(hf);
    }
    // peaks (match_template.cpp:1839-1980): integer maximum of the z-score map away from a
    // T/4 + 1 border; mask a disc of min_peak_radius; stop below the threshold or at 1000 peaks
    if (!r90 && W == ox && H == oy) {
      // the search geometry is the image's: the peak search runs on the device
      std::vector<tm_peak_info> pk(1000);
      const int found = core.FindPeaks(total, thr, T / 4 + 1, min_peak_radius, pixel_size, pk.data(), 1000);
      FILE* pf2 = std::fopen((out + "_peaks.txt").c_str(), "w");
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
      FILE* pf2 = std::fopen((out + "_peaks.txt").c_str(), "w");
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
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "match_template_hip: %s\n", e.what());
    return 1;
  }
}
