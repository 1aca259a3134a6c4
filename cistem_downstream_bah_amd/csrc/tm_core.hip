// Host side of the gfx950 template-matching inner loop and its C ABI (include/tm_hip.h).
//
// One tm_core owns one device stream and all device buffers of one search (template
// spectrum, transposed input spectrum, orientation matrices, batch work buffers, per-pixel
// state). tm_run_inner_loop walks the handle's orientations in the reference order
// (position-major, then psi; match_template.cpp:996-1001) in batches of B, so the argmax
// tie-break (earliest orientation wins) and the FP64 summation order per pixel are those of
// the CPU loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "tm_host.hpp"
#include "tm_prep_kernels.hpp"

namespace {

// Row bands per orientation of the projection normalisation (k_norm_stats / k_norm_scale blocks).
constexpr int NORM_BANDS = 8;

bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

bool image_size_compiled(int n) {
  switch (n) {
#define TM_CASE(k) case k:
    TM_IMAGE_SIZES(TM_CASE)
#undef TM_CASE
    return true;
    default: return false;
  }
}

std::string image_sizes_text() {
  std::string t;
#define TM_CASE(k) t += (t.empty() ? "" : ", ") + std::to_string(k);
  TM_IMAGE_SIZES(TM_CASE)
#undef TM_CASE
  return t;
}

enum KernelId { K_EXTRACT, K_SMALLFFT, K_NORMALIZE, K_FWD_ROWS, K_MID, K_LAST, K_COUNT };
const char* kKernelNames[K_COUNT] = {"extract", "small_fft", "normalize", "fwd_rows", "mid_columns", "last_rows_epilogue"};

}  // namespace

struct tm_core {
  int device = 0;
  hipStream_t stream = nullptr;
  bool initialised = false;
  tm_init_params prm{};
  int W = 0, H = 0, Wh = 0, T = 0, Th = 0, P = 0, Ph = 0, B = 16;
  int zst = TM_PLANE_FP32;  // storage of the inter-pass plane Z
  int column_pass = TM_COLUMNS_AUTO;  // column pass in use (TM_COLUMNS_*)
  int zt = 4;                         // column-tile width of Z
  int persistent_grid = 512;          // blocks of the persistent column pass (two per CU, a multiple of 8)
  float zscale = 1.0f;      // power of two applied to 16-bit Z values before storing
  int n_pos_local = 0, n_psi = 0, n_orient = 0;
  std::vector<float> grid, psi;      // host copies ({phi, theta} per position; psi list)
  std::vector<int2> planes;          // (size_i, defocus_i) per plane index, in call order
  int next_orientation = 0;          // first unsearched orientation of the last plane (n_orient: done)
  long long n_ccfs = 0;
  // device data
  DevBuf<cf> tpl, inpT, phaseP, phaseT, filt, slice, projT, rows, Z;
  DevBuf<float> mats;
  DevBuf<float> mip;
  DevBuf<int> idx;
  DevBuf<double> sum, sumsq;
  DevBuf<double> norm_part;  // projection normalisation: NORM_BANDS band sums per orientation
  DevBuf<unsigned long long> hist;
  std::map<int, cf*> tw;
  // timing: event pairs recorded around each kernel group on the handle's stream, resolved after
  // the call's final synchronisation (no synchronisation inside the loop)
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<std::pair<int, size_t>> ev_marks;  // (kernel id, index of the start event)
  double kernel_ms[K_COUNT] = {};
  long long kernel_launches[K_COUNT] = {};
  // graph mode: the launches of one plane captured once into a hipGraph and replayed per plane;
  // the plane's index offset lives in device memory (plane_base), written before each replay
  bool use_graph = false, capturing = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  DevBuf<int> plane_base;
  int* h_plane_base = nullptr;  // pinned host staging of the offset
  // device preparation of a micrograph (tm_set_micrograph)
  DevBuf<float> pr_src, pr_pad, pr_st, pr_cx, pr_cy;
  DevBuf<cf> pr_spec;
  DevBuf<double> pr_part, pr_acc, pr_pc, pr_ps;
  // bumped whenever the search buffers are freed (tm_init, tm_destroy): a group graph that captured
  // this handle's buffer addresses under an older epoch is stale
  unsigned long long epoch = 0;

  void drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
  }

  ~tm_core() {
    drop_graph();
    if (h_plane_base) (void)hipHostFree(h_plane_base);
    for (auto& kv : tw) (void)hipFree(kv.second);
    for (auto& e : ev_pool) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  hipEvent_t next_event() {
    if (ev_used == ev_pool.size()) {
      hipEvent_t e;
      TM_HIP(hipEventCreate(&e));
      ev_pool.push_back(e);
    }
    return ev_pool[ev_used++];
  }

  void resolve_timing() {
    for (auto& m : ev_marks) {
      float ms = 0.f;
      TM_HIP(hipEventElapsedTime(&ms, ev_pool[m.second], ev_pool[m.second + 1]));
      kernel_ms[m.first] += ms;
      kernel_launches[m.first] += 1;
    }
    ev_marks.clear();
    ev_used = 0;
  }

  const cf* twiddle(int n) {
    auto it = tw.find(n);
    if (it != tw.end()) return it->second;
    auto host = twiddles(n);
    cf* d = nullptr;
    device_alloc((void**)&d, sizeof(cf) * n);
    tw[n] = d;
    TM_HIP(hipMemcpy(d, host.data(), sizeof(cf) * n, hipMemcpyHostToDevice));
    return d;
  }

  // Frees the search's device buffers (not the stream), as before a fresh tm_init.
  void release_buffers() {
    drop_graph();
    ++epoch;
    for (DevBuf<cf>* b : {&tpl, &inpT, &phaseP, &phaseT, &filt, &slice, &projT, &rows, &Z}) b->release();
    mats.release();
    norm_part.release();
    mip.release();
    idx.release();
    sum.release();
    sumsq.release();
    hist.release();
    for (auto& kv : tw) (void)hipFree(kv.second);
    tw.clear();
  }
};

namespace {

template <int HH, int KC, int NTC>
void launch_mid_t(tm_core* h, int nb) {
  if constexpr (NTC == 0) {
    throw TmError{TM_ERR_UNSUPPORTED, "no generic column pass for height " + std::to_string(HH)};
  } else {
  const int ntiles = (h->Wh + KC - 1) / KC;
  const int grid = ((ntiles + 7) / 8) * 8;
  // orientations spread over grid rows until about 2048 blocks are in flight
  const int rows_y = std::max(1, std::min(nb, (2048 + grid - 1) / grid));
  const dim3 g(grid, rows_y), bl(KC * NTC);
  if (h->zst == TM_PLANE_FP16)
    hipLaunchKernelGGL((tmk::k_mid<HH, KC, NTC, tmk::ZS_F16>), g, bl, 0, h->stream, h->rows.p, h->inpT.p, h->Z.p, h->Wh,
                       h->T, nb, ntiles, h->twiddle(HH), h->zscale);
  else if (h->zst == TM_PLANE_BF16)
    hipLaunchKernelGGL((tmk::k_mid<HH, KC, NTC, tmk::ZS_BF16>), g, bl, 0, h->stream, h->rows.p, h->inpT.p, h->Z.p, h->Wh,
                       h->T, nb, ntiles, h->twiddle(HH), h->zscale);
  else
    hipLaunchKernelGGL((tmk::k_mid<HH, KC, NTC>), g, bl, 0, h->stream, h->rows.p, h->inpT.p, h->Z.p, h->Wh, h->T, nb,
                       ntiles, h->twiddle(HH), h->zscale);
  check_launch("k_mid");
  }
}

// Columns per k_mid block for image height H, and the column-tile width of the Z layout
// Z[b][tile][y][z_tile] (the handle's h->zt, set at init by column_pass_for):
//   * persistent pass at H = 4096, T <= 256 (k_mid4096p): 4-column tiles (the default; the halves
//     written by two blocks at once merge in L2) or 2-column tiles, each written whole by one
//     wavefront (faster columns, slower row reads);
//   * one block per column pair at H = 4096 (k_mid4096): two 2-column blocks fill one 4-column tile,
//     so that a row read of k_last takes 32 bytes from each cache line instead of 16;
//   * other heights (k_mid): one tile per block, tmk::mid_cols(H) columns wide, or half as wide with
//     TM_COLUMNS_NARROW (a launch-shape choice for the autotuner: more, smaller blocks per CU against
//     shorter row-pass reads).
int mid_kc(int H) { return H == 4096 ? 2 : tmk::mid_cols(H); }
int mid_tiles(const tm_core* h) { return (h->Wh + mid_kc(h->H) - 1) / mid_kc(h->H); }
// Whether the column pass places the projection rows at y = 0 (with the centring phase folded into
// the input spectrum) rather than at H/2 - T/2.
bool column_rows_at_zero(int H) { return H == 4096; }
int z_tiles(const tm_core* h) { return (h->Wh + h->zt - 1) / h->zt; }
// Whether the row pass of a handle with Z tiles zt columns wide is the direct-load k_last_dma (power-of-two
// widths from 1024, tiles at least one 16-byte load wide: two FP32 or four 16-bit elements) rather
// than the gathering k_last. The two agree to FP32 rounding, not bit for bit (k_last_dma takes its
// twiddles from factored tables).
bool row_pass_dma(const tm_core* h, int zt) {
  return h->W >= 1024 && (h->W & (h->W - 1)) == 0 && zt >= (h->zst == TM_PLANE_FP32 ? 2 : 4);
}

// The column pass of a handle: the persistent pass with 4-column tiles where it applies (H = 4096,
// T <= 256, FP32 planes) unless the caller asks for another; the Z tile width follows.
void column_pass_for(tm_core* h, int requested) {
  // 16-bit planes are read four elements per 16-byte load by the row pass, so their tiles are 4 wide:
  // the persistent pass with 2-column tiles takes FP32 planes only
  const bool persistent_ok = h->H == 4096 && h->T <= 256;
  TM_REQUIRE(requested != TM_COLUMNS_PERSISTENT || (persistent_ok && h->zst == TM_PLANE_FP32), TM_ERR_UNSUPPORTED,
             "the persistent column pass with 2-column tiles needs image height 4096, projection size <= 256 "
             "and FP32 planes");
  TM_REQUIRE(requested != TM_COLUMNS_PAIRS || h->H == 4096, TM_ERR_UNSUPPORTED,
             "the column-pair pass needs image height 4096");
  TM_REQUIRE(requested != TM_COLUMNS_PERSISTENT4 || persistent_ok, TM_ERR_UNSUPPORTED,
             "the persistent column pass needs image height 4096 and projection size <= 256");
  // narrow tiles only where the row pass stays the same kernel, so that results stay bit-identical
  TM_REQUIRE(requested != TM_COLUMNS_NARROW ||
                 (h->H != 4096 && row_pass_dma(h, mid_kc(h->H)) == row_pass_dma(h, mid_kc(h->H) / 2)),
             TM_ERR_UNSUPPORTED,
             "the narrow column tiles apply to image heights other than 4096 where they keep the row pass "
             "(not at " + std::to_string(h->W) + "x" + std::to_string(h->H) + " with this plane storage)");
  if (h->H != 4096) {
    // one column tile per block: tmk::mid_cols(H) columns, or half as many (narrow)
    h->column_pass = requested == TM_COLUMNS_NARROW ? TM_COLUMNS_NARROW : TM_COLUMNS_AUTO;
    h->zt = requested == TM_COLUMNS_NARROW ? mid_kc(h->H) / 2 : mid_kc(h->H);
  } else if (requested == TM_COLUMNS_PAIRS || !persistent_ok) {
    h->column_pass = TM_COLUMNS_PAIRS;
    h->zt = 4;
  } else if (requested == TM_COLUMNS_PERSISTENT) {
    h->column_pass = TM_COLUMNS_PERSISTENT;
    h->zt = 2;
  } else {  // TM_COLUMNS_AUTO or TM_COLUMNS_PERSISTENT4: 4-column tiles keep the row pass's reads wide
    h->column_pass = TM_COLUMNS_PERSISTENT4;
    h->zt = 4;
  }
}

template <int ZST>
void launch_mid4096(tm_core* h, int nb) {
  if (h->column_pass == TM_COLUMNS_PERSISTENT || h->column_pass == TM_COLUMNS_PERSISTENT4) {
    if (h->column_pass == TM_COLUMNS_PERSISTENT)
      hipLaunchKernelGGL((tmk::k_mid4096p<ZST, 2>), dim3(h->persistent_grid), dim3(512), 0, h->stream, h->rows.p,
                         h->inpT.p, h->Z.p, h->Wh, h->T, nb, (h->Wh + 1) / 2, h->twiddle(4096), h->zscale);
    else
      hipLaunchKernelGGL((tmk::k_mid4096p<ZST, 4>), dim3(h->persistent_grid), dim3(512), 0, h->stream, h->rows.p,
                         h->inpT.p, h->Z.p, h->Wh, h->T, nb, (h->Wh + 1) / 2, h->twiddle(4096), h->zscale);
    check_launch("k_mid4096p");
    return;
  }
  const int ntiles = mid_tiles(h);
  const int grid = ((ntiles + 7) / 8) * 8;
  if (h->T <= 256)
    hipLaunchKernelGGL((tmk::k_mid4096<2, true, 4, true, true, ZST>), dim3(grid), dim3(2 * 256), 0, h->stream,
                       h->rows.p, h->inpT.p, h->Z.p, h->Wh, h->T, nb, ntiles, h->twiddle(4096), h->zscale);
  else
    hipLaunchKernelGGL((tmk::k_mid4096<2, true, 4, true, false, ZST>), dim3(grid), dim3(2 * 256), 0, h->stream,
                       h->rows.p, h->inpT.p, h->Z.p, h->Wh, h->T, nb, ntiles, h->twiddle(4096), h->zscale);
  check_launch("k_mid4096");
}

void launch_mid(tm_core* h, int nb) {
  if (h->H == 4096) {
    switch (h->zst) {
      case TM_PLANE_FP16: launch_mid4096<tmk::ZS_F16>(h, nb); break;
      case TM_PLANE_BF16: launch_mid4096<tmk::ZS_BF16>(h, nb); break;
      default: launch_mid4096<tmk::ZS_F32>(h, nb); break;
    }
    return;
  }
  const bool narrow = h->column_pass == TM_COLUMNS_NARROW;
  switch (h->H) {
#define TM_CASE(n)                                                                      \
  case n:                                                                               \
    if (narrow) launch_mid_t<n, tmk::mid_cols(n) / 2, tmk::mid_threads(n)>(h, nb);     \
    else launch_mid_t<n, tmk::mid_cols(n), tmk::mid_threads(n)>(h, nb);                \
    break;
    TM_IMAGE_SIZES(TM_CASE)
#undef TM_CASE
    default: throw TmError{TM_ERR_UNSUPPORTED, "unsupported image height " + std::to_string(h->H)};
  }
}

// QR orientations per k_last block (tmk::last_qr). For power-of-two widths from 1024 on the
// groups are whole wavefronts and the row gather runs as double-buffered direct LDS loads; those read
// 16 bytes of one tile row at a time (two FP32 or four 16-bit elements), so planes in tiles narrower
// than that take the gathering row pass.
template <int WW, int QR>
void launch_last_t(tm_core* h, int nb, tmk::EpiParams ep) {
  ep.ztile = h->zt;
  ep.ntiles = z_tiles(h);
  ep.hstep_inv = 1.0 / ep.hstep;
  ep.zinv = 1.0f / h->zscale;  // a power of two: exact
  const dim3 g(h->H);
  if constexpr (WW >= 1024 && (WW & (WW - 1)) == 0) {
    if (row_pass_dma(h, h->zt)) {
      const dim3 bl(QR * (WW / 16));
      if (h->zst == TM_PLANE_FP16)
        hipLaunchKernelGGL((tmk::k_last_dma<WW, QR, tmk::ZS_F16>), g, bl, 0, h->stream, h->Z.p, h->H, nb, ep,
                           h->twiddle(WW), h->twiddle(WW / 2));
      else if (h->zst == TM_PLANE_BF16)
        hipLaunchKernelGGL((tmk::k_last_dma<WW, QR, tmk::ZS_BF16>), g, bl, 0, h->stream, h->Z.p, h->H, nb, ep,
                           h->twiddle(WW), h->twiddle(WW / 2));
      else
        hipLaunchKernelGGL((tmk::k_last_dma<WW, QR>), g, bl, 0, h->stream, h->Z.p, h->H, nb, ep, h->twiddle(WW),
                           h->twiddle(WW / 2));
      check_launch("k_last_dma");
      return;
    }
  }
  const dim3 bl(tmk::last_threads<WW, QR>());
  if (h->zst == TM_PLANE_FP16)
    hipLaunchKernelGGL((tmk::k_last<WW, QR, tmk::ZS_F16>), g, bl, 0, h->stream, h->Z.p, h->H, nb, ep, h->twiddle(WW),
                       h->twiddle(WW / 2));
  else if (h->zst == TM_PLANE_BF16)
    hipLaunchKernelGGL((tmk::k_last<WW, QR, tmk::ZS_BF16>), g, bl, 0, h->stream, h->Z.p, h->H, nb, ep, h->twiddle(WW),
                       h->twiddle(WW / 2));
  else
    hipLaunchKernelGGL((tmk::k_last<WW, QR>), g, bl, 0, h->stream, h->Z.p, h->H, nb, ep, h->twiddle(WW),
                       h->twiddle(WW / 2));
  check_launch("k_last");
}

void launch_last(tm_core* h, int nb, const tmk::EpiParams& ep) {
  switch (h->W) {
#define TM_CASE(n) \
  case n: launch_last_t<n, tmk::last_qr(n)>(h, nb, ep); break;
    TM_IMAGE_SIZES(TM_CASE)
#undef TM_CASE
    default: throw TmError{TM_ERR_UNSUPPORTED, "unsupported image width " + std::to_string(h->W)};
  }
}

struct Timer {
  tm_core* h;
  int k;
  size_t start = 0;
  Timer(tm_core* hh, int kk) : h(hh), k(kk) {
    if (h->timing) {
      start = h->ev_used;
      TM_HIP(hipEventRecord(h->next_event(), h->stream));
    }
  }
  // A failed record here surfaces on the next checked HIP call of the same stream.
  ~Timer() {
    if (h->timing) {
      hipEvent_t e = nullptr;
      try { e = h->next_event(); } catch (...) { return; }
      if (hipEventRecord(e, h->stream) == hipSuccess) h->ev_marks.push_back({k, start});
    }
  }
};

// ------------------------------------------------------------------ batch pipeline
// Normalised real projections for orientations [o0, o0 + nb) into projT (T x (T+2) floats each).
void run_projections(tm_core* h, int o0, int nb) {
  const int P = h->P, Ph = h->Ph, T = h->T, Th = h->Th;
  const size_t slice_bs = (size_t)P * Ph;     // complex
  const size_t proj_bs = (size_t)T * Th;      // complex (= T * (T+2) floats)
  {
    // ExtractSlice (x the swap phase, and when P == T the projection filter), then the column half
    // of the C2R (BackwardFFT)
    Timer t(h, K_EXTRACT);
    if (P != T) extract_cols(h, P, h->tpl.p, h->mats.p, o0, nb, h->phaseP.p, nullptr, h->slice.p);
    else extract_cols(h, T, h->tpl.p, h->mats.p, o0, nb, h->phaseT.p, h->filt.p, h->projT.p);
  }
  {
    Timer t(h, K_SMALLFFT);
    if (P != T) {
      // row half of the C2R, only for the T rows that ClipInto keeps (centre rows P/2 - T/2 ..)
      const int r0 = P / 2 - T / 2;
      rows_c2r(h, P, h->slice.p + (size_t)r0 * Ph, slice_bs, Ph,
               reinterpret_cast<float*>(h->slice.p) + (size_t)r0 * (P + 2), 2 * slice_bs, P + 2, T, nb);
      // ClipInto T^2 (centre) + ForwardFFT rows; columns fused with the 1/T^2 scale, filter and inverse
      rows_r2c(h, T, reinterpret_cast<const float*>(h->slice.p), 2 * slice_bs, P + 2, r0, r0, T, 0, h->projT.p,
               proj_bs, Th, T, nb, 1.f);
      const float scale = (float)(1. / (double)(float)((long long)T * T));
      cols(h, T, h->projT.p, Th, proj_bs, nb, -1, 1, h->filt.p, scale);
    }
    rows_c2r(h, T, h->projT.p, proj_bs, Th, reinterpret_cast<float*>(h->projT.p), 2 * proj_bs, T + 2, T, nb);
  }
  {
    Timer t(h, K_NORMALIZE);
    hipLaunchKernelGGL((tmk::k_norm_stats<256>), dim3(NORM_BANDS, nb), dim3(256), 0, h->stream,
                       reinterpret_cast<const float*>(h->projT.p), T, 2 * proj_bs, h->norm_part.p);
    check_launch("k_norm_stats");
    hipLaunchKernelGGL((tmk::k_norm_scale<256>), dim3(NORM_BANDS, nb), dim3(256), 0, h->stream,
                       reinterpret_cast<float*>(h->projT.p), T, 2 * proj_bs, h->norm_part.p, (long long)h->W * h->H);
    check_launch("k_norm_scale");
  }
}

// Forward row FFTs of the padded reference: projection row r sits at column W/2 - T/2 of
// row H/2 - T/2 + r (ClipIntoLargerRealSpace2D); scaled by 1/(W H) (ForwardFFT). With a 16-bit
// plane the rows are stored in the same 16-bit format, times the plane's power-of-two scale (the
// column pass removes it exactly when it widens them).
void run_forward_rows(tm_core* h, int nb) {
  Timer t(h, K_FWD_ROWS);
  const int T = h->T;
  const float scale = (float)(1. / (double)(float)((long long)h->W * h->H));
  const float* src = reinterpret_cast<const float*>(h->projT.p);
  const size_t sbs = 2 * (size_t)T * h->Th, dbs = (size_t)T * h->Wh;
  const int c0 = h->W / 2 - T / 2;
  if (h->zst == TM_PLANE_FP16)
    rows_r2c_16<tmk::ZS_F16>(h, h->W, src, sbs, T + 2, 0, 0, T, c0, h->rows.p, dbs, h->Wh, T, nb, scale * h->zscale);
  else if (h->zst == TM_PLANE_BF16)
    rows_r2c_16<tmk::ZS_BF16>(h, h->W, src, sbs, T + 2, 0, 0, T, c0, h->rows.p, dbs, h->Wh, T, nb, scale * h->zscale);
  else
    rows_r2c(h, h->W, src, sbs, T + 2, 0, 0, T, c0, h->rows.p, dbs, h->Wh, T, nb, scale);
}

void run_batch(tm_core* h, int o0, int nb, int g0) {
  run_projections(h, o0, nb);
  run_forward_rows(h, nb);
  {
    Timer t(h, K_MID);
    launch_mid(h, nb);
  }
  {
    Timer t(h, K_LAST);
    tmk::EpiParams ep{};
    ep.mip = h->mip.p;
    ep.idx = h->idx.p;
    ep.sum = h->sum.p;
    ep.sumsq = h->sumsq.p;
    ep.hist = h->hist.p;
    ep.hmin = h->prm.histogram_min_scaled;
    ep.hstep = h->prm.histogram_step_scaled;
    ep.nbins = h->prm.histogram_number_of_bins;
    ep.g0 = g0;
    ep.plane_g0 = h->capturing ? h->plane_base.p : nullptr;
    launch_last(h, nb, ep);
  }
}

void upload_filter(tm_core* h, const float* projection_filter) {
  TM_REQUIRE(projection_filter, TM_ERR_INVALID, "projection_filter is NULL");
  const size_t n = (size_t)h->T * h->Th;
  TM_HIP(hipMemcpyAsync(h->filt.p, projection_filter, n * sizeof(cf), hipMemcpyHostToDevice, h->stream));
}

// The input spectrum, transposed to [kx][ky] so that each column is contiguous. At H = 4096 the
// column pass places the projection rows at y = 0..T-1 instead of the reference's centred
// y0 = H/2 - T/2 (ClipIntoLargerRealSpace2D); the shift multiplies the reference spectrum by
// e^{-2 pi i y0 ky / H}, so conj(R) * I keeps its value with I' = I * e^{+2 pi i y0 ky / H}.
void upload_input(tm_core* h, const float* input_spectrum) {
  const int H = h->H, T = h->T;
  std::vector<cf> it((size_t)h->Wh * H);
  const cf* in = reinterpret_cast<const cf*>(input_spectrum);
  const bool shifted = column_rows_at_zero(H);
  std::vector<long double> pc(H, 1.0L), ps(H, 0.0L);
  if (shifted) {
    const long double two_pi = 2.0L * acosl(-1.0L);
    const long long y0 = H / 2 - T / 2;
    for (int y = 0; y < H; ++y) {
      const long double a = two_pi * (long double)((y0 * y) % H) / (long double)H;
      pc[y] = cosl(a);
      ps[y] = sinl(a);
    }
  }
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < h->Wh; ++x) {
      const cf v = in[(size_t)y * h->Wh + x];
      cf& o = it[(size_t)x * H + y];
      if (shifted) {
        o.x = (float)(v.x * pc[y] - v.y * ps[y]);
        o.y = (float)(v.x * ps[y] + v.y * pc[y]);
      } else {
        o = v;
      }
    }
  if (h->inpT.n != it.size()) h->inpT.alloc(it.size());
  TM_HIP(hipMemcpyAsync(h->inpT.p, it.data(), it.size() * sizeof(cf), hipMemcpyHostToDevice, h->stream));
  TM_HIP(hipStreamSynchronize(h->stream));
}

void reset_state(tm_core* h) {
  const size_t n = (size_t)h->W * h->H;
  hipLaunchKernelGGL(tmk::k_reset_state, dim3((n + 255) / 256), dim3(256), 0, h->stream, h->mip.p, h->idx.p, h->sum.p,
                     h->sumsq.p, n);
  check_launch("k_reset_state");
  hipLaunchKernelGGL(tmk::k_zero_u64, dim3(2), dim3(256), 0, h->stream, h->hist.p, 512);
  check_launch("k_zero_u64");
  h->planes.clear();
  h->next_orientation = 0;
  h->n_ccfs = 0;
}

void require_ready(tm_core* h) {
  TM_REQUIRE(h, TM_ERR_INVALID, "handle is NULL");
  TM_REQUIRE(h->initialised, TM_ERR_STATE, "tm_init has not succeeded on this handle");
  TM_HIP(hipSetDevice(h->device));
}

// ------------------------------------------------------------------ device peak search
// First maximum in row-major order of z over [edge, W - edge) x [edge, H - edge): per-block
// (value, index) partials over rows, then one block combines them (ties keep the lower index).
__device__ inline bool zbetter(float v, int i, float bv, int bi) { return v > bv || (v == bv && i < bi); }

__global__ __launch_bounds__(256) void k_zmax_part(const float* __restrict__ z, int W, int H, int edge,
                                                   float* __restrict__ pv, int* __restrict__ pi) {
  __shared__ float sv[256];
  __shared__ int si[256];
  float bv = -INFINITY;
  int bi = INT_MAX;
  const int x0 = edge, x1 = W - edge, y1 = H - edge;
  for (int y = edge + blockIdx.x; y < y1; y += gridDim.x)
    for (int x = x0 + threadIdx.x; x < x1; x += 256) {
      const int i = y * W + x;
      const float v = z[i];
      if (zbetter(v, i, bv, bi)) { bv = v; bi = i; }
    }
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k && zbetter(sv[threadIdx.x + k], si[threadIdx.x + k], sv[threadIdx.x], si[threadIdx.x])) {
      sv[threadIdx.x] = sv[threadIdx.x + k];
      si[threadIdx.x] = si[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = sv[0];
    pi[blockIdx.x] = si[0];
  }
}

// the maximum of n partials, its pixel's best orientation index: out = {value bits, pixel, index}
__global__ __launch_bounds__(256) void k_zmax_final(const float* __restrict__ pv, const int* __restrict__ pi, int n,
                                                    const int* __restrict__ idx, int* __restrict__ out) {
  __shared__ float sv[256];
  __shared__ int si[256];
  float bv = -INFINITY;
  int bi = INT_MAX;
  for (int q = threadIdx.x; q < n; q += 256)
    if (zbetter(pv[q], pi[q], bv, bi)) { bv = pv[q]; bi = pi[q]; }
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k && zbetter(sv[threadIdx.x + k], si[threadIdx.x + k], sv[threadIdx.x], si[threadIdx.x])) {
      sv[threadIdx.x] = sv[threadIdx.x + k];
      si[threadIdx.x] = si[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = __float_as_int(sv[0]);
    out[1] = si[0];
    out[2] = si[0] == INT_MAX ? -1 : idx[si[0]];
  }
}

// -FLT_MAX over the pixels of [by - r, by + r) x [bx - r, bx + r) whose squared distance is <= r2
__global__ void k_zmask(float* __restrict__ z, int W, int H, int bx, int by, int r, float r2) {
  const int x = max(bx - r, 0) + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int y = max(by - r, 0) + (int)blockIdx.y;
  if (x >= min(bx + r, W) || y >= min(by + r, H)) return;
  const float dx = float(x) - bx, dy = float(y) - by;
  if (dx * dx + dy * dy <= r2) z[(size_t)y * W + x] = -FLT_MAX;
}

// ------------------------------------------------------------------ allocation-free peak search
// The same search as above with no z-score map: each pixel's z-score is computed where it is read
// (finalize_pixel, the arithmetic of k_finalize) and the exclusion discs of the peaks found so far
// are one bit per pixel. Its working memory is the file-scope array `scratch` below: the bits, then
// the reduction partials and the result.
__global__ __launch_bounds__(256) void k_fill_u32(unsigned* __restrict__ p, size_t n, unsigned v) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = v;
}

__global__ __launch_bounds__(256) void k_zmax_part_bits(const float* __restrict__ mip, const double* __restrict__ sum,
                                                        const double* __restrict__ sumsq,
                                                        const unsigned* __restrict__ excluded, int W, int H, int edge,
                                                        float n_ccf, float sqrt_px, float* __restrict__ pv,
                                                        int* __restrict__ pi) {
  __shared__ float sv[256];
  __shared__ int si[256];
  float bv = -INFINITY;
  int bi = INT_MAX;
  const int x0 = edge, x1 = W - edge, y1 = H - edge;
  for (int y = edge + blockIdx.x; y < y1; y += gridDim.x)
    for (int x = x0 + threadIdx.x; x < x1; x += 256) {
      const int i = y * W + x;
      const float v = ((excluded[i >> 5] >> (i & 31)) & 1u)
                          ? -FLT_MAX
                          : tmk::finalize_pixel(mip[i], sum[i], sumsq[i], n_ccf, sqrt_px).z;
      if (zbetter(v, i, bv, bi)) { bv = v; bi = i; }
    }
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k && zbetter(sv[threadIdx.x + k], si[threadIdx.x + k], sv[threadIdx.x], si[threadIdx.x])) {
      sv[threadIdx.x] = sv[threadIdx.x + k];
      si[threadIdx.x] = si[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = sv[0];
    pi[blockIdx.x] = si[0];
  }
}

// k_zmask's disc, as set bits
__global__ void k_zmask_bits(unsigned* __restrict__ excluded, int W, int H, int bx, int by, int r, float r2) {
  const int x = max(bx - r, 0) + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int y = max(by - r, 0) + (int)blockIdx.y;
  if (x >= min(bx + r, W) || y >= min(by + r, H)) return;
  const float dx = float(x) - bx, dy = float(y) - by;
  if (dx * dx + dy * dy <= r2) {
    const int i = y * W + x;
    atomicOr(&excluded[i >> 5], 1u << (i & 31));
  }
}

}  // namespace

// Working memory of tm_find_peaks_noalloc (TM_NOALLOC_SCRATCH_BYTES): every device that loads the
// library holds its own copy until the process ends. Host code never passes `scratch` itself to a
// kernel (on the host it is only a stand-in for the symbol): tm_create looks up each device's
// address once, with that device current, and the entry point uses that address.
static __device__ __align__(256) unsigned char scratch[TM_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

// once per device, with that device current
void scratch_setup(int device) {
  TM_REQUIRE(device >= 0 && device < kMaxDevices, TM_ERR_INVALID, "device ordinal beyond the library's table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    throw TmError{TM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
  g_scratch[device] = p;
}

void* scratch_address(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  void* p = (device >= 0 && device < kMaxDevices) ? g_scratch[device] : nullptr;
  TM_REQUIRE(p, TM_ERR_STATE, "no scratch address for device " + std::to_string(device));
  return p;
}

// The peak loop of tm_find_peaks and tm_find_peaks_noalloc: find_max() leaves {value bits, pixel,
// orientation index} of the highest remaining z-score at res on the device; exclude() removes the
// disc of min_peak_radius around a peak (box [bx - r, bx + r) x [by - r, by + r) clipped to the
// image, one 64-thread row of blocks per image row). Returns the number of peaks written.
template <class FindMax, class Exclude>
int peak_loop(tm_core* h, float threshold, int edge, float min_peak_radius, float pixel_size, tm_peak_info* out,
              int capacity, const int* res, FindMax find_max, Exclude exclude) {
  const int W = h->W, H = h->H;
  hipStream_t st = h->stream;
  const int r = (int)min_peak_radius;
  const float r2 = powf(float(r), 2);
  const int per_plane = h->prm.n_search_positions * h->n_psi;
  int found = 0;
  while (found < capacity && W > 2 * edge && H > 2 * edge) {
    find_max();
    int hres[3];
    TM_HIP(hipMemcpyAsync(hres, res, sizeof(hres), hipMemcpyDeviceToHost, st));
    TM_HIP(hipStreamSynchronize(st));
    float v;
    std::memcpy(&v, &hres[0], sizeof(v));
    if (hres[1] < 0 || hres[1] == INT_MAX || !(v >= threshold)) break;
    const int bx = hres[1] % W, by = hres[1] / W, g = hres[2];
    float a_psi = 0.f, a_theta = 0.f, a_phi = 0.f, a_def = 0.f, a_pix = 0.f;
    if (g >= 0) {
      const int plane = g / per_plane, rem = g % per_plane;
      const int pos = rem / h->n_psi, k = rem % h->n_psi;
      a_psi = h->psi[k];
      a_theta = h->grid[2 * pos + 1];
      a_phi = h->grid[2 * pos];
      a_def = float(h->planes[plane].y) * h->prm.defocus_step;
      a_pix = float(h->planes[plane].x) * h->prm.pixel_size_step;
    }
    tm_peak_info& o = out[found];
    ++found;
    o.peak_number = o.original_peak_number = o.new_peak_number = found;
    o.x_pos = (float)bx * pixel_size;
    o.y_pos = (float)by * pixel_size;
    o.psi = a_psi;
    o.theta = a_theta;
    o.phi = a_phi;
    o.defocus = a_def;
    o.pixel_size = a_pix;
    o.peak_height = v;
    const int x_lo = std::max(bx - r, 0), x_hi = std::min(bx + r, W), y_lo = std::max(by - r, 0), y_hi = std::min(by + r, H);
    if (x_hi > x_lo && y_hi > y_lo) exclude(bx, by, r, r2, dim3((x_hi - x_lo + 63) / 64, y_hi - y_lo));
  }
  TM_HIP(hipStreamSynchronize(st));
  return found;
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

const char* tm_last_error(void) { return g_err.c_str(); }

int tm_device_count(int* count) {
  return guarded([&] {
    TM_REQUIRE(count, TM_ERR_INVALID, "count is NULL");
    int n = 0;
    TM_HIP(hipGetDeviceCount(&n));
    *count = n;
  });
}

int tm_create(int device_ordinal, tm_core** out) {
  return guarded([&] {
    TM_REQUIRE(out, TM_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int n = 0;
    TM_HIP(hipGetDeviceCount(&n));
    TM_REQUIRE(device_ordinal >= 0 && device_ordinal < n, TM_ERR_INVALID,
               "device ordinal " + std::to_string(device_ordinal) + " out of range (" + std::to_string(n) + " devices)");
    TM_HIP(hipSetDevice(device_ordinal));
    scratch_setup(device_ordinal);
    auto* h = new tm_core();
    h->device = device_ordinal;
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete h;
      throw TmError{TM_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)};
    }
    *out = h;
  });
}

int tm_destroy(tm_core* h) {
  return guarded([&] {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    delete h;
  });
}

int tm_init(tm_core* h, const tm_init_params* p, const float* template_spectrum, const float* input_spectrum,
            const float* search_positions, const float* psi_list) {
  return guarded([&] {
    TM_REQUIRE(h && p, TM_ERR_INVALID, "NULL handle or params");
    TM_REQUIRE(template_spectrum && input_spectrum && search_positions && psi_list, TM_ERR_INVALID,
               "NULL input array");
    TM_HIP(hipSetDevice(h->device));
    h->initialised = false;
    TM_REQUIRE(p->device_memory_limit >= 0 && p->orientation_limit >= 0, TM_ERR_INVALID,
               "device_memory_limit and orientation_limit must be >= 0");
    TM_HIP(hipStreamSynchronize(h->stream));
    h->release_buffers();
    // every allocation below counts against device_memory_limit; on failure the buffers are freed
    ScopedAllocLimit limit(p->device_memory_limit);
    struct FreeOnError {
      tm_core* h;
      bool ok = false;
      ~FreeOnError() { if (!ok) h->release_buffers(); }
    } free_on_error{h};
    const int P = p->template_size, T = p->projection_size, W = p->image_x, H = p->image_y;
    TM_REQUIRE(image_size_compiled(W) && image_size_compiled(H), TM_ERR_UNSUPPORTED,
               "image size " + std::to_string(W) + "x" + std::to_string(H) +
                   " not supported (sides " + image_sizes_text() + ")");
    TM_REQUIRE(pow2(T) && T >= 32 && T <= 1024 && pow2(P) && P >= T && P <= 1024, TM_ERR_UNSUPPORTED,
               "template/projection sizes " + std::to_string(P) + "/" + std::to_string(T) +
                   " not supported (powers of two, 32 <= T <= P <= 1024)");
    TM_REQUIRE(T <= W && T <= H, TM_ERR_INVALID, "projection larger than the image");
    TM_REQUIRE(p->n_search_positions > 0 && p->n_psi > 0, TM_ERR_INVALID, "empty search grid");
    TM_REQUIRE(p->first_search_position >= 0 && p->last_search_position < p->n_search_positions &&
                   p->first_search_position <= p->last_search_position,
               TM_ERR_INVALID, "search position range out of bounds");
    TM_REQUIRE(p->histogram_number_of_bins > 0 && p->histogram_number_of_bins <= 512, TM_ERR_INVALID,
               "histogram_number_of_bins must be in 1..512");
    TM_REQUIRE(p->histogram_step_scaled > 0.0, TM_ERR_INVALID, "histogram_step_scaled must be > 0");
    h->drop_graph();
    h->prm = *p;
    h->W = W; h->H = H; h->Wh = W / 2 + 1;
    h->T = T; h->Th = T / 2 + 1;
    h->P = P; h->Ph = P / 2 + 1;
    h->B = p->batch_size > 0 ? p->batch_size : 64;
    TM_REQUIRE(p->plane_storage == TM_PLANE_FP32 || p->plane_storage == TM_PLANE_FP16 ||
                   p->plane_storage == TM_PLANE_BF16, TM_ERR_INVALID, "plane_storage must be 0, 1 or 2");
    h->zst = p->plane_storage;
    // 16-bit values: the plane's elements are ~ 1 / sqrt(W^2 H) (correlations ~ 1 / sqrt(W H) after
    // the row transform of W points); scale them to ~1/4, leaving FP16 headroom above and below
    h->zscale = h->zst == TM_PLANE_FP32 ? 1.0f
                                       : ldexpf(1.0f, (int)lround(0.5 * std::log2((double)W * W * H)) - 2);
    TM_REQUIRE(h->B <= 256, TM_ERR_INVALID, "batch_size must be <= 256");
    TM_REQUIRE(p->column_pass >= TM_COLUMNS_AUTO && p->column_pass <= TM_COLUMNS_NARROW, TM_ERR_INVALID,
               "column_pass must be 0, 1, 2, 3 or 4");
    column_pass_for(h, p->column_pass);
    int cus = 0;
    TM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
    h->persistent_grid = std::max(8, (2 * cus) / 8 * 8);
    h->n_pos_local = p->last_search_position - p->first_search_position + 1;
    h->n_psi = p->n_psi;
    long long n_orient = (long long)h->n_pos_local * h->n_psi;
    if (p->orientation_limit > 0) n_orient = std::min<long long>(n_orient, p->orientation_limit);
    TM_REQUIRE(n_orient < (1LL << 30), TM_ERR_INVALID, "too many orientations for one handle");
    h->n_orient = (int)n_orient;
    h->n_pos_local = (int)((n_orient + h->n_psi - 1) / h->n_psi);  // positions with at least one orientation
    h->grid.assign(search_positions, search_positions + 2 * (size_t)p->n_search_positions);
    h->psi.assign(psi_list, psi_list + p->n_psi);

    // orientation matrices in reference loop order: position-major, psi inner
    std::vector<float> m((size_t)h->n_orient * 9);
    for (int o = 0; o < h->n_orient; ++o) {
      const int pos = p->first_search_position + o / h->n_psi, k = o % h->n_psi;
      euler_matrix(h->grid[2 * pos], h->grid[2 * pos + 1], h->psi[k], &m[(size_t)o * 9]);
    }
    h->mats.alloc(m.size());
    TM_HIP(hipMemcpy(h->mats.p, m.data(), m.size() * sizeof(float), hipMemcpyHostToDevice));

    // template spectrum
    const size_t ntpl = (size_t)P * P * h->Ph;
    h->tpl.alloc(ntpl);
    TM_HIP(hipMemcpy(h->tpl.p, template_spectrum, ntpl * sizeof(cf), hipMemcpyHostToDevice));

    upload_input(h, input_spectrum);

    // phase tables and filter
    auto pp = swap_phase(P);
    h->phaseP.alloc(pp.size());
    TM_HIP(hipMemcpy(h->phaseP.p, pp.data(), pp.size() * sizeof(cf), hipMemcpyHostToDevice));
    auto pt = swap_phase(T);
    h->phaseT.alloc(pt.size());
    TM_HIP(hipMemcpy(h->phaseT.p, pt.data(), pt.size() * sizeof(cf), hipMemcpyHostToDevice));
    h->filt.alloc((size_t)T * h->Th);

    // batch work buffers
    const int B = h->B;
    h->slice.alloc((size_t)B * P * h->Ph);
    h->projT.alloc((size_t)B * T * h->Th);
    const size_t row_elems = (size_t)B * T * h->Wh;  // forward rows: stored like the plane
    h->rows.alloc(h->zst == TM_PLANE_FP32 ? row_elems : (row_elems + 1) / 2);
    const size_t z_elems = (size_t)B * z_tiles(h) * H * h->zt;
    h->Z.alloc(h->zst == TM_PLANE_FP32 ? z_elems : (z_elems + 1) / 2);  // cf units; 16-bit: 4 bytes each

    // state
    const size_t npx = (size_t)W * H;
    h->mip.alloc(npx);
    h->idx.alloc(npx);
    h->sum.alloc(npx);
    h->sumsq.alloc(npx);
    h->hist.alloc(512);
    h->norm_part.alloc((size_t)B * NORM_BANDS * 3);
    for (int k : {P, P / 2, T, T / 2, W, W / 2, H}) h->twiddle(k);
    reset_state(h);
    TM_HIP(hipStreamSynchronize(h->stream));
    free_on_error.ok = true;
    h->initialised = true;
  });
}

int tm_set_template(tm_core* h, const float* template_spectrum) {
  return guarded([&] {
    require_ready(h);
    TM_REQUIRE(template_spectrum, TM_ERR_INVALID, "template_spectrum is NULL");
    TM_HIP(hipMemcpyAsync(h->tpl.p, template_spectrum, h->tpl.n * sizeof(cf), hipMemcpyHostToDevice, h->stream));
    TM_HIP(hipStreamSynchronize(h->stream));
  });
}

int tm_set_input(tm_core* h, const float* input_spectrum) {
  return guarded([&] {
    require_ready(h);
    TM_REQUIRE(input_spectrum, TM_ERR_INVALID, "input_spectrum is NULL");
    TM_HIP(hipStreamSynchronize(h->stream));
    upload_input(h, input_spectrum);
    reset_state(h);
    TM_HIP(hipStreamSynchronize(h->stream));
  });
}

int tm_set_micrograph(tm_core* h, const float* micrograph, float* curve_x, float* curve_y) {
  return guarded([&] {
    require_ready(h);
    TM_REQUIRE(micrograph, TM_ERR_INVALID, "micrograph is NULL");
    const int W = h->W, H = h->H, nh = W / 2 + 1;
    const int n = int((W / 2.0 + 1.0) * sqrtf(2.0) + 1.0);
    const long long npx = (long long)W * H, nspec = (long long)nh * H;
    hipStream_t st = h->stream;
    TM_HIP(hipStreamSynchronize(st));
    if (h->pr_src.n != (size_t)npx) h->pr_src.alloc(npx);
    if (h->pr_pad.n != (size_t)(W + 2) * H) h->pr_pad.alloc((size_t)(W + 2) * H);
    if (h->pr_spec.n != (size_t)nspec) h->pr_spec.alloc(nspec);
    if (!h->pr_part.p) h->pr_part.alloc(2 * 1024);
    if (!h->pr_st.p) h->pr_st.alloc(3);
    if (h->pr_acc.n != (size_t)2 * n) h->pr_acc.alloc(2 * (size_t)n);
    if (h->pr_cx.n != (size_t)n) { h->pr_cx.alloc(n); h->pr_cy.alloc(n); }
    TM_HIP(hipMemcpyAsync(h->pr_src.p, micrograph, npx * sizeof(float), hipMemcpyHostToDevice, st));
    // ReplaceOutliersWithMean(5) (image.cpp:986-1100)
    const long long chunk = std::max<long long>(256, (npx + 1023) / 1024);
    const int np = (int)((npx + chunk - 1) / chunk);
    hipLaunchKernelGGL(tmp::k_moments<256>, dim3(np), dim3(256), 0, st, h->pr_src.p, npx, chunk, h->pr_part.p);
    check_launch("k_moments");
    hipLaunchKernelGGL(tmp::k_outlier_bounds, dim3(1), dim3(1), 0, st, h->pr_part.p, np, npx, h->pr_st.p);
    check_launch("k_outlier_bounds");
    const unsigned g = (unsigned)std::min<long long>(65535, ((long long)(W + 2) * H + 255) / 256);
    hipLaunchKernelGGL(tmp::k_outlier_pad, dim3(g), dim3(256), 0, st, h->pr_src.p, W, H, h->pr_st.p, h->pr_pad.p);
    check_launch("k_outlier_pad");
    // ForwardFFT (scaled 1/N), SwapRealSpaceQuadrants, ZeroCentralPixel (match_template.cpp:703-709)
    rows_r2c(h, W, h->pr_pad.p, (size_t)(W + 2) * H, W + 2, 0, 0, W, 0, h->pr_spec.p, nspec, nh, H, 1, 1.0f);
    cols_line(h, H, h->pr_spec.p, nh, nspec, 1, -1);
    const float inv_n = 1. / (float)((long)W * H);
    const unsigned gs = (unsigned)std::min<long long>(65535, (nspec + 255) / 256);
    hipLaunchKernelGGL(tmp::k_scale_zero_dc, dim3(gs), dim3(256), 0, st, h->pr_spec.p, nspec, inv_n);
    check_launch("k_scale_zero_dc");
    hipLaunchKernelGGL(tmp::k_swap_phase_2d, dim3(gs), dim3(256), 0, st, h->pr_spec.p, W, H);
    check_launch("k_swap_phase_2d");
    // whitening curve (Curve::SetupXAxis(0, 0.5 sqrt 2, n), match_template.cpp:711-716)
    std::vector<float> x(n);
    const float lo = 0.0f, hi = 0.5 * sqrtf(2.0);
    for (int c = 0; c < n; ++c) x[c] = c * (hi - lo) / float(n - 1) + lo;
    TM_HIP(hipMemcpyAsync(h->pr_cx.p, x.data(), n * sizeof(float), hipMemcpyHostToDevice, st));
    TM_HIP(hipMemsetAsync(h->pr_acc.p, 0, 2 * (size_t)n * sizeof(double), st));
    tmp::power_curve(st, h->pr_spec.p, W, H, h->pr_cx.p, n, h->pr_acc.p, h->pr_acc.p + n, gs);
    check_launch("k_power_curve");
    hipLaunchKernelGGL(tmp::k_curve_finish, dim3(1), dim3(256), 0, st, h->pr_acc.p, h->pr_acc.p + n, n, h->pr_cy.p);
    check_launch("k_curve_finish");
    // ApplyCurveFilter, ZeroCentralPixel, 1 / sqrt(ReturnSumOfSquares()) (match_template.cpp:717-724)
    hipLaunchKernelGGL(tmp::k_apply_curve, dim3(gs), dim3(256), 0, st, h->pr_spec.p, W, H, h->pr_cx.p, h->pr_cy.p, n);
    check_launch("k_apply_curve");
    const long long schunk = std::max<long long>(256, (nspec + 1023) / 1024);
    const int nsp = (int)((nspec + schunk - 1) / schunk);
    hipLaunchKernelGGL(tmp::k_spectrum_ssq<256>, dim3(nsp), dim3(256), 0, st, h->pr_spec.p, W, H, schunk,
                       h->pr_part.p);
    check_launch("k_spectrum_ssq");
    hipLaunchKernelGGL(tmp::k_normalise_spectrum, dim3(gs), dim3(256), 0, st, h->pr_spec.p, nspec, h->pr_part.p, nsp,
                       1);
    check_launch("k_normalise_spectrum");
    // the handle's input spectrum: transposed, with the centring phase at H = 4096 (as upload_input)
    const double* pc = nullptr;
    const double* ps = nullptr;
    if (column_rows_at_zero(H)) {
      if (h->pr_pc.n != (size_t)H) { h->pr_pc.alloc(H); h->pr_ps.alloc(H); }
      std::vector<double> c(H), sn(H);
      const long double two_pi = 2.0L * acosl(-1.0L);
      const long long y0 = H / 2 - h->T / 2;
      for (int y = 0; y < H; ++y) {
        const long double a = two_pi * (long double)((y0 * y) % H) / (long double)H;
        c[y] = (double)cosl(a);
        sn[y] = (double)sinl(a);
      }
      TM_HIP(hipMemcpyAsync(h->pr_pc.p, c.data(), H * sizeof(double), hipMemcpyHostToDevice, st));
      TM_HIP(hipMemcpyAsync(h->pr_ps.p, sn.data(), H * sizeof(double), hipMemcpyHostToDevice, st));
      pc = h->pr_pc.p;
      ps = h->pr_ps.p;
    }
    hipLaunchKernelGGL(tmp::k_transpose_phase, dim3(gs), dim3(256), 0, st, h->pr_spec.p, nh, H, pc, ps, h->inpT.p);
    check_launch("k_transpose_phase");
    if (curve_x) TM_HIP(hipMemcpyAsync(curve_x, h->pr_cx.p, n * sizeof(float), hipMemcpyDeviceToHost, st));
    if (curve_y) TM_HIP(hipMemcpyAsync(curve_y, h->pr_cy.p, n * sizeof(float), hipMemcpyDeviceToHost, st));
    reset_state(h);
    TM_HIP(hipStreamSynchronize(st));
  });
}

int tm_reset(tm_core* h) {
  return guarded([&] {
    require_ready(h);
    reset_state(h);
    TM_HIP(hipStreamSynchronize(h->stream));
  });
}

int tm_run_inner_loop(tm_core* h, const float* projection_filter, int size_i, int defocus_i, tm_progress_fn progress,
                      void* user, long long* ccfs) {
  return guarded([&] {
    require_ready(h);
    upload_filter(h, projection_filter);
    const int plane = (int)h->planes.size();
    // global orientation index: (plane * n_positions + pos) * n_psi + psi over the whole grid, so
    // handles that search different position ranges produce comparable, loop-ordered indices
    const long long per_plane = (long long)h->prm.n_search_positions * h->n_psi;
    TM_REQUIRE((long long)(plane + 1) * per_plane < (1LL << 31), TM_ERR_INVALID,
               "orientation index space exhausted (too many planes for one handle)");
    const long long first = (long long)h->prm.first_search_position * h->n_psi;
    h->planes.push_back(make_int2(size_i, defocus_i));
    for (int k = 0; k < K_COUNT; ++k) { h->kernel_ms[k] = 0.0; h->kernel_launches[k] = 0; }
    h->ev_marks.clear();
    h->ev_used = 0;
    long long done = 0;
    if (h->use_graph && !h->timing) {
      // graph mode: the plane's offset goes to device memory; the batches (indices relative to the
      // plane) are captured once and replayed for every plane
      *h->h_plane_base = (int)(plane * per_plane);
      TM_HIP(hipMemcpyAsync(h->plane_base.p, h->h_plane_base, sizeof(int), hipMemcpyHostToDevice, h->stream));
      if (!h->graph_exec) {
        TM_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        h->capturing = true;
        try {
          for (int o0 = 0; o0 < h->n_orient; o0 += h->B)
            run_batch(h, o0, std::min(h->B, h->n_orient - o0), (int)(first + o0));
        } catch (...) {
          hipGraph_t g = nullptr;
          (void)hipStreamEndCapture(h->stream, &g);
          if (g) (void)hipGraphDestroy(g);
          h->capturing = false;
          throw;
        }
        h->capturing = false;
        TM_HIP(hipStreamEndCapture(h->stream, &h->graph));
        TM_HIP(hipGraphInstantiate(&h->graph_exec, h->graph, nullptr, nullptr, 0));
      }
      TM_HIP(hipGraphLaunch(h->graph_exec, h->stream));
      done = h->n_orient;
      if (progress) {
        TM_HIP(hipStreamSynchronize(h->stream));
        progress(done, user);
      }
    } else {
      for (int o0 = 0; o0 < h->n_orient; o0 += h->B) {
        const int nb = std::min(h->B, h->n_orient - o0);
        run_batch(h, o0, nb, (int)(plane * per_plane + first + o0));
        done += nb;
        if (progress) {
          TM_HIP(hipStreamSynchronize(h->stream));
          progress(done, user);
        }
      }
    }
    TM_HIP(hipStreamSynchronize(h->stream));
    if (h->timing) h->resolve_timing();
    h->n_ccfs += done;
    h->next_orientation = h->n_orient;
    if (ccfs) *ccfs = done;
  });
}

int tm_run_partial(tm_core* h, const float* projection_filter, int size_i, int defocus_i, int first_orientation,
                   int n_orientations, long long* ccfs) {
  return guarded([&] {
    require_ready(h);
    TM_REQUIRE(first_orientation >= 0 && n_orientations >= 0 &&
                   (long long)first_orientation + n_orientations <= h->n_orient,
               TM_ERR_INVALID, "orientation range outside the handle's " + std::to_string(h->n_orient));
    if (first_orientation == 0) {
      const long long per_plane = (long long)h->prm.n_search_positions * h->n_psi;
      TM_REQUIRE((long long)(h->planes.size() + 1) * per_plane < (1LL << 31), TM_ERR_INVALID,
                 "orientation index space exhausted (too many planes for one handle)");
      TM_REQUIRE(h->planes.empty() || h->next_orientation == h->n_orient, TM_ERR_STATE,
                 "the previous plane is not finished (next orientation " + std::to_string(h->next_orientation) + ")");
      h->planes.push_back(make_int2(size_i, defocus_i));
      h->next_orientation = 0;
    } else {
      TM_REQUIRE(!h->planes.empty() && h->planes.back().x == size_i && h->planes.back().y == defocus_i,
                 TM_ERR_STATE, "a continued range must belong to the current plane");
      TM_REQUIRE(first_orientation == h->next_orientation, TM_ERR_STATE,
                 "the current plane continues at orientation " + std::to_string(h->next_orientation));
    }
    upload_filter(h, projection_filter);
    const int plane = (int)h->planes.size() - 1;
    const long long per_plane = (long long)h->prm.n_search_positions * h->n_psi;
    const long long first = (long long)h->prm.first_search_position * h->n_psi;
    for (int k = 0; k < K_COUNT; ++k) { h->kernel_ms[k] = 0.0; h->kernel_launches[k] = 0; }
    h->ev_marks.clear();
    h->ev_used = 0;
    const int end = first_orientation + n_orientations;
    for (int o0 = first_orientation; o0 < end; o0 += h->B) {
      const int nb = std::min(h->B, end - o0);
      run_batch(h, o0, nb, (int)(plane * per_plane + first + o0));
    }
    TM_HIP(hipStreamSynchronize(h->stream));
    if (h->timing) h->resolve_timing();
    h->n_ccfs += n_orientations;
    h->next_orientation = end;
    if (ccfs) *ccfs = n_orientations;
  });
}

namespace {
// Serialised search state: header, planes, then the maps and the histogram.
struct StateHeader {
  char magic[8];
  int W, H, n_orient, n_psi, n_search_positions, first_search_position, n_planes, next_orientation;
  long long n_ccfs;
};
constexpr char kStateMagic[8] = {'T', 'M', 'S', 'T', 'A', 'T', 'E', '1'};

long long state_bytes(const tm_core* h, size_t n_planes) {
  const long long npx = (long long)h->W * h->H;
  return (long long)sizeof(StateHeader) + (long long)n_planes * 2 * sizeof(int) + npx * (4 + 4 + 8 + 8) + 512 * 8;
}
}  // namespace

int tm_state_size(tm_core* h, long long* bytes) {
  return guarded([&] {
    require_ready(h);
    TM_REQUIRE(bytes, TM_ERR_INVALID, "bytes is NULL");
    *bytes = state_bytes(h, h->planes.size());
  });
}

int tm_export_state(tm_core* h, void* buffer, long long bytes) {
  return guarded([&] {
    require_ready(h);
    TM_REQUIRE(buffer, TM_ERR_INVALID, "buffer is NULL");
    TM_REQUIRE(bytes >= state_bytes(h, h->planes.size()), TM_ERR_INVALID, "buffer smaller than tm_state_size");
    StateHeader hd{};
    std::memcpy(hd.magic, kStateMagic, sizeof(hd.magic));
    hd.W = h->W; hd.H = h->H; hd.n_orient = h->n_orient; hd.n_psi = h->n_psi;
    hd.n_search_positions = h->prm.n_search_positions;
    hd.first_search_position = h->prm.first_search_position;
    hd.n_planes = (int)h->planes.size();
    hd.next_orientation = h->next_orientation;
    hd.n_ccfs = h->n_ccfs;
    char* o = static_cast<char*>(buffer);
    std::memcpy(o, &hd, sizeof(hd));
    o += sizeof(hd);
    for (const int2& pl : h->planes) {
      const int v[2] = {pl.x, pl.y};
      std::memcpy(o, v, sizeof(v));
      o += sizeof(v);
    }
    const size_t npx = (size_t)h->W * h->H;
    hipStream_t st = h->stream;
    TM_HIP(hipMemcpyAsync(o, h->mip.p, npx * 4, hipMemcpyDeviceToHost, st));
    o += npx * 4;
    TM_HIP(hipMemcpyAsync(o, h->idx.p, npx * 4, hipMemcpyDeviceToHost, st));
    o += npx * 4;
    TM_HIP(hipMemcpyAsync(o, h->sum.p, npx * 8, hipMemcpyDeviceToHost, st));
    o += npx * 8;
    TM_HIP(hipMemcpyAsync(o, h->sumsq.p, npx * 8, hipMemcpyDeviceToHost, st));
    o += npx * 8;
    TM_HIP(hipMemcpyAsync(o, h->hist.p, 512 * 8, hipMemcpyDeviceToHost, st));
    TM_HIP(hipStreamSynchronize(st));
  });
}

int tm_import_state(tm_core* h, const void* buffer, long long bytes) {
  return guarded([&] {
    require_ready(h);
    TM_REQUIRE(buffer && bytes >= (long long)sizeof(StateHeader), TM_ERR_INVALID, "buffer is NULL or too small");
    StateHeader hd;
    const char* in = static_cast<const char*>(buffer);
    std::memcpy(&hd, in, sizeof(hd));
    in += sizeof(hd);
    TM_REQUIRE(std::memcmp(hd.magic, kStateMagic, sizeof(hd.magic)) == 0, TM_ERR_INVALID, "not a search state");
    TM_REQUIRE(hd.W == h->W && hd.H == h->H && hd.n_orient == h->n_orient && hd.n_psi == h->n_psi &&
                   hd.n_search_positions == h->prm.n_search_positions &&
                   hd.first_search_position == h->prm.first_search_position,
               TM_ERR_INVALID, "the state belongs to a different search (image, grid, psi or range)");
    TM_REQUIRE(hd.n_planes >= 0 && hd.next_orientation >= 0 && hd.next_orientation <= hd.n_orient &&
                   bytes >= state_bytes(h, (size_t)hd.n_planes),
               TM_ERR_INVALID, "truncated or inconsistent search state");
    std::vector<int2> planes(hd.n_planes);
    for (int k = 0; k < hd.n_planes; ++k) {
      int v[2];
      std::memcpy(v, in, sizeof(v));
      in += sizeof(v);
      planes[k] = make_int2(v[0], v[1]);
    }
    const size_t npx = (size_t)h->W * h->H;
    hipStream_t st = h->stream;
    TM_HIP(hipStreamSynchronize(st));
    TM_HIP(hipMemcpyAsync(h->mip.p, in, npx * 4, hipMemcpyHostToDevice, st));
    in += npx * 4;
    TM_HIP(hipMemcpyAsync(h->idx.p, in, npx * 4, hipMemcpyHostToDevice, st));
    in += npx * 4;
    TM_HIP(hipMemcpyAsync(h->sum.p, in, npx * 8, hipMemcpyHostToDevice, st));
    in += npx * 8;
    TM_HIP(hipMemcpyAsync(h->sumsq.p, in, npx * 8, hipMemcpyHostToDevice, st));
    in += npx * 8;
    TM_HIP(hipMemcpyAsync(h->hist.p, in, 512 * 8, hipMemcpyHostToDevice, st));
    TM_HIP(hipStreamSynchronize(st));
    h->planes = planes;
    h->next_orientation = hd.next_orientation;
    h->n_ccfs = hd.n_ccfs;
  });
}

int tm_get_results(tm_core* h, float* mip, float* psi, float* theta, float* phi, float* defocus, float* pixel_size,
                   double* sum, double* sum_of_squares, int* best_index, long long* histogram, long long* n_ccfs) {
  return guarded([&] {
    require_ready(h);
    const int W = h->W, H = h->H, RS = W + 2;
    const size_t npx = (size_t)W * H;
    std::vector<float> m(npx);
    std::vector<int> ix(npx);
    TM_HIP(hipMemcpyAsync(m.data(), h->mip.p, npx * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    TM_HIP(hipMemcpyAsync(ix.data(), h->idx.p, npx * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    std::vector<double> s, q;
    if (sum) {
      s.resize(npx);
      TM_HIP(hipMemcpyAsync(s.data(), h->sum.p, npx * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    }
    if (sum_of_squares) {
      q.resize(npx);
      TM_HIP(hipMemcpyAsync(q.data(), h->sumsq.p, npx * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    }
    std::vector<unsigned long long> hh(512);
    TM_HIP(hipMemcpyAsync(hh.data(), h->hist.p, 512 * sizeof(unsigned long long), hipMemcpyDeviceToHost, h->stream));
    TM_HIP(hipStreamSynchronize(h->stream));
    const int per_plane = h->prm.n_search_positions * h->n_psi;
    for (int y = 0; y < H; ++y) {
      for (int x = 0; x < RS; ++x) {
        const size_t o = (size_t)y * RS + x;
        if (x >= W) {
          if (mip) mip[o] = 0.f;
          if (psi) psi[o] = 0.f;
          if (theta) theta[o] = 0.f;
          if (phi) phi[o] = 0.f;
          if (defocus) defocus[o] = 0.f;
          if (pixel_size) pixel_size[o] = 0.f;
          if (sum) sum[o] = 0.0;
          if (sum_of_squares) sum_of_squares[o] = 0.0;
          if (best_index) best_index[o] = -1;
          continue;
        }
        const size_t i = (size_t)y * W + x;
        const int g = ix[i];
        if (mip) mip[o] = m[i];
        if (best_index) best_index[o] = g;
        if (sum) sum[o] = s[i];
        if (sum_of_squares) sum_of_squares[o] = q[i];
        float a_psi = 0.f, a_theta = 0.f, a_phi = 0.f, a_def = 0.f, a_pix = 0.f;
        if (g >= 0) {
          const int plane = g / per_plane;
          const int rem = g % per_plane;
          const int pos = rem / h->n_psi;
          const int k = rem % h->n_psi;
          a_psi = h->psi[k];
          a_theta = h->grid[2 * pos + 1];
          a_phi = h->grid[2 * pos];
          a_def = float(h->planes[plane].y) * h->prm.defocus_step;
          a_pix = float(h->planes[plane].x) * h->prm.pixel_size_step;
        }
        if (psi) psi[o] = a_psi;
        if (theta) theta[o] = a_theta;
        if (phi) phi[o] = a_phi;
        if (defocus) defocus[o] = a_def;
        if (pixel_size) pixel_size[o] = a_pix;
      }
    }
    if (histogram)
      for (int b = 0; b < h->prm.histogram_number_of_bins; ++b) histogram[b] = (long long)hh[b];
    if (n_ccfs) *n_ccfs = h->n_ccfs;
  });
}

int tm_merge(tm_core* dst, tm_core* src) {
  return guarded([&] {
    require_ready(dst);
    require_ready(src);
    TM_REQUIRE(dst != src, TM_ERR_INVALID, "cannot merge a handle into itself");
    TM_REQUIRE(dst->device == src->device, TM_ERR_INVALID, "handles on different devices");
    TM_REQUIRE(dst->W == src->W && dst->H == src->H && dst->n_psi == src->n_psi &&
                   dst->prm.n_search_positions == src->prm.n_search_positions &&
                   dst->prm.histogram_number_of_bins == src->prm.histogram_number_of_bins &&
                   dst->prm.histogram_min_scaled == src->prm.histogram_min_scaled &&
                   dst->prm.histogram_step_scaled == src->prm.histogram_step_scaled,
               TM_ERR_INVALID, "handles describe different searches (image, grid, psi or histogram)");
    TM_REQUIRE(dst->planes.size() == src->planes.size(), TM_ERR_INVALID, "handles ran different plane sequences");
    for (size_t k = 0; k < dst->planes.size(); ++k)
      TM_REQUIRE(dst->planes[k].x == src->planes[k].x && dst->planes[k].y == src->planes[k].y, TM_ERR_INVALID,
                 "handles ran different plane sequences");
    TM_HIP(hipSetDevice(dst->device));
    TM_HIP(hipStreamSynchronize(src->stream));
    const size_t n = (size_t)dst->W * dst->H;
    hipLaunchKernelGGL(tmk::k_merge, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, dst->stream, dst->mip.p,
                       dst->idx.p, dst->sum.p, dst->sumsq.p, src->mip.p, src->idx.p, src->sum.p, src->sumsq.p, n);
    check_launch("k_merge");
    hipLaunchKernelGGL(tmk::k_add_u64, dim3(2), dim3(256), 0, dst->stream, dst->hist.p, src->hist.p, 512);
    check_launch("k_add_u64");
    TM_HIP(hipStreamSynchronize(dst->stream));
    dst->n_ccfs += src->n_ccfs;
  });
}

int tm_find_peaks(tm_core* h, long long total_ccfs, float threshold, int edge, float min_peak_radius, float pixel_size,
                  tm_peak_info* out, int capacity, int* n_peaks) {
  return guarded([&] {
    require_ready(h);
    TM_REQUIRE(total_ccfs > 0 && n_peaks && (out || capacity == 0) && capacity >= 0 && edge >= 0, TM_ERR_INVALID,
               "invalid arguments");
    const int W = h->W, H = h->H;
    const size_t npx = (size_t)W * H;
    hipStream_t st = h->stream;
    DevBuf<float> fin, pv;
    DevBuf<int> pi, res;
    fin.alloc(4 * npx);
    const int nparts = std::max(1, std::min(TM_NOALLOC_PARTS, H - 2 * edge));
    pv.alloc(nparts);
    pi.alloc(nparts);
    res.alloc(3);
    const double sqrt_px = sqrt((double)(W * H));
    hipLaunchKernelGGL(tmk::k_finalize, dim3((npx + 255) / 256), dim3(256), 0, st, h->mip.p, h->sum.p, h->sumsq.p, npx,
                       (float)total_ccfs, (float)sqrt_px, fin.p, fin.p + npx, fin.p + 2 * npx, fin.p + 3 * npx);
    check_launch("k_finalize");
    float* z = fin.p + npx;
    *n_peaks = peak_loop(
        h, threshold, edge, min_peak_radius, pixel_size, out, capacity, res.p,
        [&] {
          hipLaunchKernelGGL(k_zmax_part, dim3(nparts), dim3(256), 0, st, z, W, H, edge, pv.p, pi.p);
          check_launch("k_zmax_part");
          hipLaunchKernelGGL(k_zmax_final, dim3(1), dim3(256), 0, st, pv.p, pi.p, nparts, h->idx.p, res.p);
          check_launch("k_zmax_final");
        },
        [&](int bx, int by, int r, float r2, dim3 grid) {
          hipLaunchKernelGGL(k_zmask, grid, dim3(64), 0, st, z, W, H, bx, by, r, r2);
          check_launch("k_zmask");
        });
  });
}

int tm_find_peaks_noalloc(tm_core* h, long long total_ccfs, float threshold, int edge, float min_peak_radius,
                          float pixel_size, tm_peak_info* out, int capacity, int* n_peaks) {
  return guarded([&] {
    require_ready(h);
    TM_REQUIRE(total_ccfs > 0 && n_peaks && (out || capacity == 0) && capacity >= 0 && edge >= 0, TM_ERR_INVALID,
               "invalid arguments");
    const int W = h->W, H = h->H;
    const long long npx = (long long)W * H;
    TM_REQUIRE(npx <= TM_NOALLOC_MAX_PIXELS, TM_ERR_INVALID,
               "search image of " + std::to_string(W) + " x " + std::to_string(H) +
                   " pixels exceeds the allocation-free peak search's limit of " +
                   std::to_string(TM_NOALLOC_MAX_PIXELS) + " pixels");
    unsigned* excluded = static_cast<unsigned*>(scratch_address(h->device));
    float* pv = reinterpret_cast<float*>(excluded + TM_NOALLOC_MAX_PIXELS / 32);
    int* pi = reinterpret_cast<int*>(pv + TM_NOALLOC_PARTS);
    int* res = pi + TM_NOALLOC_PARTS;
    hipStream_t st = h->stream;
    const int nparts = std::max(1, std::min(TM_NOALLOC_PARTS, H - 2 * edge));
    const size_t words = (size_t)(npx + 31) / 32;
    // the bits of the previous call on this device are still there: clear this image's share
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)std::min<size_t>((words + 255) / 256, 2048)), dim3(256), 0, st,
                       excluded, words, 0u);
    check_launch("k_fill_u32");
    const float n_ccf = (float)total_ccfs, sqrt_px = (float)sqrt((double)(W * H));
    *n_peaks = peak_loop(
        h, threshold, edge, min_peak_radius, pixel_size, out, capacity, res,
        [&] {
          hipLaunchKernelGGL(k_zmax_part_bits, dim3(nparts), dim3(256), 0, st, h->mip.p, h->sum.p, h->sumsq.p,
                             excluded, W, H, edge, n_ccf, sqrt_px, pv, pi);
          check_launch("k_zmax_part_bits");
          hipLaunchKernelGGL(k_zmax_final, dim3(1), dim3(256), 0, st, pv, pi, nparts, h->idx.p, res);
          check_launch("k_zmax_final");
        },
        [&](int bx, int by, int r, float r2, dim3 grid) {
          hipLaunchKernelGGL(k_zmask_bits, grid, dim3(64), 0, st, excluded, W, H, bx, by, r, r2);
          check_launch("k_zmask_bits");
        });
  });
}

int tm_finalize(tm_core* h, long long total_ccfs, float* mip_scaled, float* zscore, float* average, float* std_dev) {
  return guarded([&] {
    require_ready(h);
    TM_REQUIRE(total_ccfs > 0, TM_ERR_INVALID, "total_ccfs must be > 0");
    const int W = h->W, H = h->H, RS = W + 2;
    const size_t npx = (size_t)W * H;
    DevBuf<float> out;
    out.alloc(4 * npx);
    const double sqrt_px = sqrt((double)(W * H));
    hipLaunchKernelGGL(tmk::k_finalize, dim3((npx + 255) / 256), dim3(256), 0, h->stream, h->mip.p, h->sum.p,
                       h->sumsq.p, npx, (float)total_ccfs, (float)sqrt_px, out.p, out.p + npx, out.p + 2 * npx,
                       out.p + 3 * npx);
    check_launch("k_finalize");
    std::vector<float> host(4 * npx);
    TM_HIP(hipMemcpyAsync(host.data(), out.p, 4 * npx * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    TM_HIP(hipStreamSynchronize(h->stream));
    float* dst[4] = {mip_scaled, zscore, average, std_dev};
    for (int k = 0; k < 4; ++k) {
      if (!dst[k]) continue;
      for (int y = 0; y < H; ++y) {
        std::memcpy(dst[k] + (size_t)y * RS, host.data() + k * npx + (size_t)y * W, W * sizeof(float));
        dst[k][(size_t)y * RS + W] = 0.f;
        dst[k][(size_t)y * RS + W + 1] = 0.f;
      }
    }
  });
}

int tm_debug_projection(tm_core* h, const float* projection_filter, int search_position, int psi_index, float* out) {
  return guarded([&] {
    require_ready(h);
    TM_REQUIRE(out, TM_ERR_INVALID, "out is NULL");
    const int pl = search_position - h->prm.first_search_position;
    TM_REQUIRE(pl >= 0 && pl < h->n_pos_local && psi_index >= 0 && psi_index < h->n_psi &&
                   pl * h->n_psi + psi_index < h->n_orient,
               TM_ERR_INVALID, "orientation outside the handle's range");
    upload_filter(h, projection_filter);
    run_projections(h, pl * h->n_psi + psi_index, 1);
    TM_HIP(hipMemcpyAsync(out, h->projT.p, (size_t)h->T * (h->T + 2) * sizeof(float), hipMemcpyDeviceToHost,
                          h->stream));
    TM_HIP(hipStreamSynchronize(h->stream));
  });
}

int tm_debug_ccf(tm_core* h, const float* projection_filter, int search_position, int psi_index, float* out) {
  return guarded([&] {
    require_ready(h);
    TM_REQUIRE(out, TM_ERR_INVALID, "out is NULL");
    const int pl = search_position - h->prm.first_search_position;
    TM_REQUIRE(pl >= 0 && pl < h->n_pos_local && psi_index >= 0 && psi_index < h->n_psi &&
                   pl * h->n_psi + psi_index < h->n_orient,
               TM_ERR_INVALID, "orientation outside the handle's range");
    upload_filter(h, projection_filter);
    run_projections(h, pl * h->n_psi + psi_index, 1);
    run_forward_rows(h, 1);
    launch_mid(h, 1);
    // C2R of one plane with a scratch state: MIP of a single CCF is the CCF itself
    const int W = h->W, H = h->H;
    const size_t npx = (size_t)W * H;
    DevBuf<float> m;
    DevBuf<int> ix;
    DevBuf<double> s, q;
    DevBuf<unsigned long long> hs;
    m.alloc(npx); ix.alloc(npx); s.alloc(npx); q.alloc(npx); hs.alloc(512);
    hipLaunchKernelGGL(tmk::k_reset_state, dim3((npx + 255) / 256), dim3(256), 0, h->stream, m.p, ix.p, s.p, q.p, npx);
    check_launch("k_reset_state");
    hipLaunchKernelGGL(tmk::k_zero_u64, dim3(2), dim3(256), 0, h->stream, hs.p, 512);
    tmk::EpiParams ep{};
    ep.mip = m.p; ep.idx = ix.p; ep.sum = s.p; ep.sumsq = q.p; ep.hist = hs.p;
    ep.hmin = h->prm.histogram_min_scaled;
    ep.hstep = h->prm.histogram_step_scaled;
    ep.nbins = h->prm.histogram_number_of_bins;
    ep.g0 = 0;
    launch_last(h, 1, ep);
    std::vector<float> host(npx);
    TM_HIP(hipMemcpyAsync(host.data(), m.p, npx * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    TM_HIP(hipStreamSynchronize(h->stream));
    for (int y = 0; y < H; ++y) {
      std::memcpy(out + (size_t)y * (W + 2), host.data() + (size_t)y * W, W * sizeof(float));
      out[(size_t)y * (W + 2) + W] = 0.f;
      out[(size_t)y * (W + 2) + W + 1] = 0.f;
    }
  });
}

int tm_set_graph(tm_core* h, int enabled) {
  return guarded([&] {
    TM_REQUIRE(h, TM_ERR_INVALID, "handle is NULL");
    TM_HIP(hipSetDevice(h->device));
    TM_HIP(hipStreamSynchronize(h->stream));
    h->drop_graph();
    h->use_graph = enabled != 0;
    if (h->use_graph && !h->plane_base.p) {
      h->plane_base.alloc(1);
      TM_HIP(hipHostMalloc((void**)&h->h_plane_base, sizeof(int), hipHostMallocDefault));
    }
  });
}

int tm_get_launch_shape(tm_core* h, int* batch_size, int* column_pass, int* z_tile) {
  return guarded([&] {
    require_ready(h);
    if (batch_size) *batch_size = h->B;
    if (column_pass) *column_pass = h->column_pass;
    if (z_tile) *z_tile = h->zt;
  });
}

int tm_set_timing(tm_core* h, int enabled) {
  return guarded([&] {
    TM_REQUIRE(h, TM_ERR_INVALID, "handle is NULL");
    h->timing = enabled != 0;
  });
}

int tm_get_timing(tm_core* h, const char** names, double* ms, long long* launches, int cap) {
  if (!h) { g_err = "handle is NULL"; return TM_ERR_INVALID; }
  int n = 0;
  for (int k = 0; k < K_COUNT && n < cap; ++k, ++n) {
    if (names) names[n] = kKernelNames[k];
    if (ms) ms[n] = h->kernel_ms[k];
    if (launches) launches[n] = h->kernel_launches[k];
  }
  return n;
}

// ------------------------------------------------------------------ groups of handles
}  // extern "C"

// Several handles (typically one micrograph each, of any sizes) whose planes run as one hipGraph. The
// handles are dealt to `lanes` chains (largest work first, each to the least-loaded chain); a chain
// runs its handles one after another, each in its own stream order, and the chains are independent
// branches of the graph, so that `lanes` searches share the GPU at a time. Captured at the first
// tm_group_run and replayed after; a handle re-initialised since (its buffers moved) makes the next
// run capture again.
struct tm_group {
  int device = 0;
  int lanes = 1;
  std::vector<tm_core*> hs;
  std::vector<unsigned long long> epochs;
  std::vector<hipEvent_t> ready, join;  // per handle: inputs uploaded (outside the graph); end of its work
  hipEvent_t fork = nullptr;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  void drop() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~tm_group() {
    drop();
    for (auto e : ready) if (e) (void)hipEventDestroy(e);
    for (auto e : join) if (e) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {
// The chains of a group: handle indices dealt, by decreasing work (pixels x orientations per plane),
// to the chain with the least work so far.
std::vector<std::vector<int>> group_chains(const tm_group* g) {
  const int n = (int)g->hs.size();
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  auto work = [&](int i) { return (double)g->hs[i]->W * g->hs[i]->H * g->hs[i]->n_orient; };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work(a) > work(b); });
  std::vector<std::vector<int>> chains(std::min(g->lanes, n));
  std::vector<double> load(chains.size(), 0.0);
  for (int i : order) {
    const size_t l = std::min_element(load.begin(), load.end()) - load.begin();
    chains[l].push_back(i);
    load[l] += work(i);
  }
  return chains;
}

// Captures one plane of every handle of the group: the group's stream forks to the first handle of
// each chain, each handle's stream waits for the previous handle of its chain, and the last handle of
// each chain joins the group's stream again.
void capture_group(tm_group* g) {
  const auto chains = group_chains(g);
  const size_t n = g->hs.size();
  std::vector<char> entered(n, 0), joined(n, 0);
  TM_HIP(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
  try {
    TM_HIP(hipEventRecord(g->fork, g->stream));
    for (const auto& chain : chains) {
      hipEvent_t prev = g->fork;
      for (int i : chain) {
        tm_core* h = g->hs[i];
        TM_HIP(hipStreamWaitEvent(h->stream, prev, 0));
        entered[i] = 1;
        h->capturing = true;
        const long long first = (long long)h->prm.first_search_position * h->n_psi;
        for (int o0 = 0; o0 < h->n_orient; o0 += h->B)
          run_batch(h, o0, std::min(h->B, h->n_orient - o0), (int)(first + o0));
        h->capturing = false;
        TM_HIP(hipEventRecord(g->join[i], h->stream));
        prev = g->join[i];
      }
      TM_HIP(hipStreamWaitEvent(g->stream, prev, 0));
      for (int i : chain) joined[i] = 1;
    }
  } catch (...) {
    // join every stream that entered the capture, so that ending it returns them all to normal
    for (size_t i = 0; i < n; ++i) {
      g->hs[i]->capturing = false;
      if (entered[i] && !joined[i] && hipEventRecord(g->join[i], g->hs[i]->stream) == hipSuccess)
        (void)hipStreamWaitEvent(g->stream, g->join[i], 0);
    }
    hipGraph_t dead = nullptr;
    (void)hipStreamEndCapture(g->stream, &dead);
    if (dead) (void)hipGraphDestroy(dead);
    throw;
  }
  TM_HIP(hipStreamEndCapture(g->stream, &g->graph));
  TM_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  for (size_t i = 0; i < n; ++i) g->epochs[i] = g->hs[i]->epoch;
}
}  // namespace

extern "C" {

int tm_group_create(tm_core* const* handles, int n, int lanes, tm_group** out) {
  return guarded([&] {
    TM_REQUIRE(out, TM_ERR_INVALID, "out is NULL");
    *out = nullptr;
    TM_REQUIRE(handles && n > 0, TM_ERR_INVALID, "a group needs at least one handle");
    TM_REQUIRE(lanes >= 0, TM_ERR_INVALID, "lanes must be >= 0");
    for (int i = 0; i < n; ++i) {
      TM_REQUIRE(handles[i], TM_ERR_INVALID, "handle " + std::to_string(i) + " is NULL");
      TM_REQUIRE(handles[i]->device == handles[0]->device, TM_ERR_INVALID, "the handles of a group share one device");
      for (int j = 0; j < i; ++j)
        TM_REQUIRE(handles[j] != handles[i], TM_ERR_INVALID, "a handle appears twice in the group");
    }
    TM_HIP(hipSetDevice(handles[0]->device));
    auto g = std::make_unique<tm_group>();
    g->device = handles[0]->device;
    g->lanes = std::min(n, lanes > 0 ? lanes : 2);
    g->hs.assign(handles, handles + n);
    g->epochs.assign(n, ~0ULL);
    g->ready.assign(n, nullptr);
    g->join.assign(n, nullptr);
    TM_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    TM_HIP(hipEventCreateWithFlags(&g->fork, hipEventDisableTiming));
    for (int i = 0; i < n; ++i) {
      TM_HIP(hipEventCreateWithFlags(&g->ready[i], hipEventDisableTiming));
      TM_HIP(hipEventCreateWithFlags(&g->join[i], hipEventDisableTiming));
    }
    *out = g.release();
  });
}

int tm_group_destroy(tm_group* g) {
  return guarded([&] {
    if (!g) return;
    (void)hipSetDevice(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    delete g;
  });
}

int tm_group_run(tm_group* g, const float* const* projection_filters, const int* size_i, const int* defocus_i,
                 long long* ccfs) {
  return guarded([&] {
    TM_REQUIRE(g, TM_ERR_INVALID, "group is NULL");
    TM_REQUIRE(projection_filters && size_i && defocus_i, TM_ERR_INVALID, "NULL argument");
    TM_HIP(hipSetDevice(g->device));
    const size_t n = g->hs.size();
    bool stale = g->exec == nullptr;
    for (size_t i = 0; i < n; ++i) {
      tm_core* h = g->hs[i];
      require_ready(h);
      TM_REQUIRE(projection_filters[i], TM_ERR_INVALID, "projection filter " + std::to_string(i) + " is NULL");
      TM_REQUIRE(!h->timing, TM_ERR_STATE, "kernel timing is on for handle " + std::to_string(i));
      TM_REQUIRE(h->planes.empty() || h->next_orientation == h->n_orient, TM_ERR_STATE,
                 "handle " + std::to_string(i) + " has an unfinished plane");
      const long long per_plane = (long long)h->prm.n_search_positions * h->n_psi;
      TM_REQUIRE((long long)(h->planes.size() + 1) * per_plane < (1LL << 31), TM_ERR_INVALID,
                 "orientation index space exhausted (too many planes for handle " + std::to_string(i) + ")");
      if (g->epochs[i] != h->epoch) stale = true;
    }
    if (stale) g->drop();
    // each handle's plane offset and filter, on its own stream, before the graph starts
    for (size_t i = 0; i < n; ++i) {
      tm_core* h = g->hs[i];
      if (!h->plane_base.p) {
        h->plane_base.alloc(1);
        TM_HIP(hipHostMalloc((void**)&h->h_plane_base, sizeof(int), hipHostMallocDefault));
      }
      const long long per_plane = (long long)h->prm.n_search_positions * h->n_psi;
      *h->h_plane_base = (int)((long long)h->planes.size() * per_plane);
      TM_HIP(hipMemcpyAsync(h->plane_base.p, h->h_plane_base, sizeof(int), hipMemcpyHostToDevice, h->stream));
      upload_filter(h, projection_filters[i]);
      TM_HIP(hipEventRecord(g->ready[i], h->stream));
      TM_HIP(hipStreamWaitEvent(g->stream, g->ready[i], 0));
    }
    if (!g->exec) capture_group(g);
    TM_HIP(hipGraphLaunch(g->exec, g->stream));
    TM_HIP(hipStreamSynchronize(g->stream));
    for (size_t i = 0; i < n; ++i) {
      tm_core* h = g->hs[i];
      h->planes.push_back(make_int2(size_i[i], defocus_i[i]));
      h->n_ccfs += h->n_orient;
      h->next_orientation = h->n_orient;
      if (ccfs) ccfs[i] = h->n_orient;
    }
  });
}

int tm_autotune(int device_ordinal, const tm_init_params* params, const float* template_spectrum,
                const float* input_spectrum, const float* search_positions, const float* psi_list,
                const float* projection_filter, const int* batch_candidates, int n_batch, const int* column_candidates,
                int n_column, int sample_orientations, int* best_batch, int* best_column, double* us_per_ccf) {
  return guarded([&] {
    TM_REQUIRE(params && template_spectrum && input_spectrum && search_positions && psi_list && projection_filter &&
                   batch_candidates && column_candidates && best_batch && best_column,
               TM_ERR_INVALID, "NULL argument");
    TM_REQUIRE(n_batch > 0 && n_column > 0 && sample_orientations > 0, TM_ERR_INVALID,
               "n_batch, n_column and sample_orientations must be positive");
    for (int c = 0; c < n_batch; ++c)
      TM_REQUIRE(batch_candidates[c] >= 1 && batch_candidates[c] <= 256, TM_ERR_INVALID,
                 "batch candidates must be in 1..256");
    for (int c = 0; c < n_column; ++c)
      TM_REQUIRE(column_candidates[c] >= TM_COLUMNS_AUTO && column_candidates[c] <= TM_COLUMNS_NARROW,
                 TM_ERR_INVALID, "column-pass candidates must be TM_COLUMNS_*");
    TM_REQUIRE(params->n_psi > 0 && params->first_search_position >= 0 &&
                   params->first_search_position <= params->last_search_position,
               TM_ERR_INVALID, "empty or invalid search position range");
    // the sample: the first orientations of the range
    tm_init_params p = *params;
    const int npos = (sample_orientations + p.n_psi - 1) / p.n_psi;
    p.last_search_position = std::min(params->last_search_position, params->first_search_position + npos - 1);
    p.orientation_limit = sample_orientations;
    auto ok = [](int rc) {
      if (rc != TM_OK) throw TmError{rc, g_err};
    };
    double best = 0.0;
    int bb = 0, bc = -1;
    for (int ci = 0; ci < n_column; ++ci) {
      for (int bi = 0; bi < n_batch; ++bi) {
        p.batch_size = batch_candidates[bi];
        p.column_pass = column_candidates[ci];
        double* slot = us_per_ccf ? us_per_ccf + (size_t)ci * n_batch + bi : nullptr;
        tm_core* h = nullptr;
        ok(tm_create(device_ordinal, &h));
        struct Drop {
          tm_core* h;
          ~Drop() { tm_destroy(h); }
        } drop{h};
        const int rc = tm_init(h, &p, template_spectrum, input_spectrum, search_positions, psi_list);
        if (rc == TM_ERR_UNSUPPORTED) {  // a column pass this search cannot use
          if (slot) *slot = -1.0;
          continue;
        }
        ok(rc);
        ok(tm_run_inner_loop(h, projection_filter, 0, 0, nullptr, nullptr, nullptr));  // warm-up plane
        long long n = 0;
        const auto t0 = std::chrono::steady_clock::now();
        ok(tm_run_inner_loop(h, projection_filter, 0, 1, nullptr, nullptr, &n));
        const double us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
        const double per = us / (double)std::max(1LL, n);
        if (slot) *slot = per;
        if (bc < 0 || per < best) {
          best = per;
          bb = batch_candidates[bi];
          bc = column_candidates[ci];
        }
      }
    }
    TM_REQUIRE(bc >= 0, TM_ERR_UNSUPPORTED, "no candidate column pass applies to this search");
    *best_batch = bb;
    *best_column = bc;
  });
}

int tm_autotune_batch(int device_ordinal, const tm_init_params* params, const float* template_spectrum,
                      const float* input_spectrum, const float* search_positions, const float* psi_list,
                      const float* projection_filter, const int* candidates, int n_candidates, int sample_orientations,
                      int* best_batch, double* us_per_ccf) {
  int column = params ? params->column_pass : 0, best_column = 0;
  return tm_autotune(device_ordinal, params, template_spectrum, input_spectrum, search_positions, psi_list,
                     projection_filter, candidates, n_candidates, &column, 1, sample_orientations, best_batch,
                     &best_column, us_per_ccf);
}

}  // extern "C"
