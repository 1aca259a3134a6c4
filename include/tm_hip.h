/*
 * C ABI of the MI355X (gfx950) template-matching inner loop.
 *
 * It replaces the reference's GPU class `TemplateMatchingCore`
 * (src/gpu/TemplateMatchingCore.h:55-179) and the pieces of `DeviceManager`
 * (src/gpu/DeviceManager.h:1-28) it uses, behind plain pointers and sizes.
 *
 * Layouts (the reference's `Image` memory, src/core/image.cpp:4761-4846):
 *   - Fourier arrays are FFTW half-complex, interleaved (re, im) float pairs:
 *       2-D  (W/2+1) * H complex, x fastest;   3-D  (P/2+1) * P * P complex, x fastest.
 *   - Real maps use the padded layout: rows of W + 2 floats (W even), logical pixels first.
 * Every call returns TM_OK (0) or a negative TM_ERR_* code; tm_last_error() describes the last
 * failure on the calling thread. Nothing in the library exits the process. One handle per
 * host thread; handles share no state.
 */
#ifndef TM_HIP_H
#define TM_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define TM_OK 0
#define TM_ERR_INVALID (-1)
#define TM_ERR_HIP (-2)
#define TM_ERR_UNSUPPORTED (-3)
#define TM_ERR_STATE (-4)

typedef struct tm_core tm_core;

/* Progress callback: number of cross-correlation maps (CCFs) finished so far in this call.
 * Replaces the per-CCF ProgressBar / JobResult updates (TemplateMatchingCore.cu:324-339). */
typedef void (*tm_progress_fn)(long long ccfs_done, void* user);

typedef struct tm_init_params {
  int template_size;          /* P: edge of the padded 3-D template (cube)                         */
  int projection_size;        /* T: edge of current_projection (defines the projection size)        */
  int image_x, image_y;       /* W, H: size of the (factorised, possibly rotated) input image        */
  int n_search_positions;     /* number of {phi, theta} rows in search_positions                      */
  int n_psi;                  /* number of in-plane angles in psi_list                               */
  int first_search_position;  /* inclusive range of search positions handled by this handle          */
  int last_search_position;
  double histogram_min_scaled;   /* match_template.cpp:633-644                                       */
  double histogram_step_scaled;
  int histogram_number_of_bins;  /* <= 512 */
  int max_padding;               /* kept for interface parity; the CPU path it mirrors does not trim */
  float defocus_step;            /* best_defocus map stores defocus_i * defocus_step                 */
  float pixel_size_step;         /* best_pixel_size map stores size_i * pixel_size_step              */
  int batch_size;                /* orientations per launch; 0 selects the default (64)              */
  int plane_storage;             /* storage of the plane between the column and row transforms:
                                    TM_PLANE_FP32 (0, the reference's CPU precision), TM_PLANE_FP16
                                    or TM_PLANE_BF16 (half the plane traffic; every compiled size).
                                    The reference's GPU path keeps its input image and padded
                                    reference in FP16 (TemplateMatchingCore.cu:163-164); here the
                                    MIP, sums and histogram stay FP32/FP64 whatever the plane.      */
  int orientation_limit;         /* 0: every orientation of the position range; N > 0: only the first
                                    N in loop order (position-major, psi inner), the last position
                                    possibly cut short (a search truncated to its first N
                                    orientations). No counterpart in the reference.                */
  int column_pass;               /* kernel of the column transforms at image height 4096 (no
                                    counterpart in the reference; results are bit-identical):
                                    TM_COLUMNS_AUTO (0): the persistent pass with 4-column Z tiles
                                    where it applies (projection size <= 256), else one block per
                                    column pair;
                                    TM_COLUMNS_PAIRS (1): one block per column pair;
                                    TM_COLUMNS_PERSISTENT (2): the persistent pass with 2-column Z
                                    tiles (FP32 planes only); TM_COLUMNS_PERSISTENT4 (3): the
                                    persistent pass with the 4-column tiles of TM_COLUMNS_PAIRS
                                    (both TM_ERR_UNSUPPORTED where they do not apply).
                                    At other heights: TM_COLUMNS_AUTO, one block per column tile of
                                    the size's compiled width, or TM_COLUMNS_NARROW (4), tiles half
                                    as wide (TM_ERR_UNSUPPORTED at height 4096, and where the narrow
                                    tiles would change the row-pass kernel: one-column tiles at
                                    widths 1024, 2048 and 4096 with FP32 planes, tiles under four
                                    columns there with 16-bit planes).                             */
  long long device_memory_limit; /* 0: none; else tm_init fails with TM_ERR_HIP, freeing what it
                                    allocated and leaving the handle uninitialised, when its device
                                    allocations would exceed this many bytes, as when hipMalloc runs
                                    out of memory (the reference exits there, defines.h:75-80).     */
} tm_init_params;

#define TM_PLANE_FP32 0
#define TM_PLANE_FP16 1
#define TM_PLANE_BF16 2

#define TM_COLUMNS_AUTO 0
#define TM_COLUMNS_PAIRS 1
#define TM_COLUMNS_PERSISTENT 2
#define TM_COLUMNS_PERSISTENT4 3
#define TM_COLUMNS_NARROW 4

/* Number of visible HIP devices. Replaces DeviceManager::Init's device query
 * (src/gpu/DeviceManager.cu:27-124). */
int tm_device_count(int* count);

/* Creates a handle bound to one device. Replaces `new TemplateMatchingCore` plus
 * DeviceManager::SetGpu (match_template.cpp:821-822, DeviceManager.cu:126-140). */
int tm_create(int device_ordinal, tm_core** out);

/* Frees every device buffer of the handle. Replaces TemplateMatchingCore::~TemplateMatchingCore
 * (TemplateMatchingCore.cu:27-38, which frees nothing). */
int tm_destroy(tm_core* h);

/* Copies the search inputs to the device and allocates the working set.
 * Replaces TemplateMatchingCore::Init (TemplateMatchingCore.h:147-171, .cu:49-138):
 *   template_spectrum  <- template_reconstruction (Fourier, quadrant-swapped, DC 0)
 *   input_spectrum     <- input_image (whitened, normalised, quadrant-swapped spectrum)
 *   search_positions   <- global_euler_search.list_of_search_parameters ({phi, theta} per row)
 *   psi_list           <- the in-plane angles of the inclusive float loop psi_start..psi_max
 * The accumulated maps are reset. */
int tm_init(tm_core* h, const tm_init_params* params, const float* template_spectrum, const float* input_spectrum,
            const float* search_positions, const float* psi_list);

/* Searches all orientations of the handle's position range for one (pixel size, defocus) plane
 * and folds the result into the accumulated maps (max with the earliest orientation winning
 * ties, FP64 sums, histogram). Replaces TemplateMatchingCore::RunInnerLoop
 * (TemplateMatchingCore.h:173, .cu:140-368) together with the host merge that follows it
 * (match_template.cpp:918-985). projection_filter is T x (T/2+1) complex (CTF x whitening).
 * ccfs (may be NULL) receives the number of CCFs computed by this call. */
int tm_run_inner_loop(tm_core* h, const float* projection_filter, int size_i, int defocus_i, tm_progress_fn progress,
                      void* user, long long* ccfs);

/* Copies the accumulated maps to caller arrays in the padded layout ((W+2) * H; padding columns
 * are written as 0). Any pointer may be NULL. Replaces the reads of d_max_intensity_projection,
 * d_best_psi/theta/phi, d_sum3, d_sumSq3 and Histogram::CopyToHostAndAdd
 * (match_template.cpp:922-984, Histogram.cu:188-203). best_index receives the global orientation
 * index ((plane * n_search_positions + search_position) * n_psi + psi_index, plane counting this
 * handle's tm_run_inner_loop calls), -1 where unset. */
int tm_get_results(tm_core* h, float* mip, float* psi, float* theta, float* phi, float* defocus, float* pixel_size,
                   double* sum, double* sum_of_squares, int* best_index, long long* histogram, long long* n_ccfs);

/* Replaces the template spectrum (same P, layout as in tm_init) and keeps the accumulated maps, for
 * the next pixel-size plane: the reference copies each plane's template into its TemplateMatchingCore
 * (match_template.cpp:837-845, 883-886). */
int tm_set_template(tm_core* h, const float* template_spectrum);

/* Clears the accumulated maps (MIP to -FLT_MAX, sums and histogram to 0). */
int tm_reset(tm_core* h);

/* Merges the results of src into dst, as AggregatedTemplateResult::AddResult merges the results of
 * separate jobs (match_template.cpp:2015-2103): per pixel the larger MIP wins with its orientation
 * (equal values keep the lower global orientation index, i.e. the one the reference's sequential
 * loop meets first), sums and histogram counts add, and dst's CCF count grows by src's. Both
 * handles must be on one device, describe the same search (image size, grid, psi list, histogram)
 * and have run the same sequence of planes; typically they cover different position ranges.
 * Merging ranges that partition a search gives the single-handle result. src is unchanged. */
int tm_merge(tm_core* dst, tm_core* src);

/* Final scaling and z-score on the device (match_template.cpp:1221-1243, 1300-1310):
 * mean = sum/N, sd = sqrt(sumsq/N - mean^2) * sqrt(W*H), mean *= sqrt(W*H),
 * mip_scaled = MIP * sqrt(W*H), z = (mip_scaled - mean) / sd (0 where sd <= 0).
 * Outputs are padded-layout float maps; any pointer may be NULL. */
int tm_finalize(tm_core* h, long long total_ccfs, float* mip_scaled, float* zscore, float* average, float* std_dev);

/* Test hooks with no counterpart in the reference: the normalised real-space projection
 * (T x (T+2) floats) and one cross-correlation plane ((W+2) x H floats) for a single orientation,
 * computed by the same kernels as tm_run_inner_loop (match_template.cpp:1004-1088). */
int tm_debug_projection(tm_core* h, const float* projection_filter, int search_position, int psi_index, float* out);
int tm_debug_ccf(tm_core* h, const float* projection_filter, int search_position, int psi_index, float* out);

/* Per hot-kernel stage over the last tm_run_inner_loop call, measured with HIP events on the
 * handle's stream when timing is enabled: total duration in ms (ms) and number of launches
 * (launches); the average per launch is ms / launches. names/ms/launches hold up to cap entries;
 * returns the number of entries written. No counterpart in the reference (diagnostics). */
int tm_set_timing(tm_core* h, int enabled);
int tm_get_timing(tm_core* h, const char** names, double* ms, long long* launches, int cap);

/* Replaces the input spectrum (same image size; layout as in tm_init) and resets the maps, keeping
 * the template, the grid, the work buffers and a captured graph: one handle searches many
 * micrographs of one size with one template. No counterpart in the reference (one image per
 * TemplateMatchingCore). */
int tm_set_input(tm_core* h, const float* input_spectrum);

/* The input preparation of match_template.cpp:703-724 on the device, for a micrograph already at
 * the handle's image size W x H (after the size policy; dense, W x H floats): ReplaceOutliersWithMean(5),
 * ForwardFFT, SwapRealSpaceQuadrants, the whitening curve and filter, normalisation; the result
 * replaces the input spectrum (as tm_set_input does, maps reset) and curve_x/curve_y (may be NULL)
 * receive the whitening curve for the projection filters (int((W/2+1) sqrt 2 + 1) points). The
 * curve's bins are summed in FP64. Micrographs that need padding or rotation go through
 * tm_prepare_input_image and tm_set_input. No counterpart in the reference. */
int tm_set_micrograph(tm_core* h, const float* micrograph, float* curve_x, float* curve_y);

/* One found peak, with the fields of TemplateMatchFoundPeakInfo in its order
 * (src/core/template_matching.h:4-18): positions in Angstroms (pixel coordinate x pixel size, as
 * match_template.cpp:1883-1884 reports them), angles in degrees, the defocus and pixel-size offsets
 * of the best plane, peak_height the z-score. peak_number, original_peak_number and new_peak_number
 * are the 1-based order in which the peak was found. */
typedef struct tm_peak_info {
  int peak_number;
  float x_pos;
  float y_pos;
  float psi;
  float theta;
  float phi;
  float defocus;
  float pixel_size;
  float peak_height;
  int original_peak_number;
  int new_peak_number;
} tm_peak_info;

/* The peak search of match_template.cpp:1839-1980 on the device, over the z-score map of the
 * search geometry (W x H, which is the micrograph's when the size policy neither pads nor rotates
 * it): the finalised z-score (as tm_finalize), then repeatedly the first maximum in row-major order
 * over [edge, W - edge) x [edge, H - edge), stopping below threshold or at capacity, each peak masking
 * the pixels of [y - r, y + r) x [x - r, x + r) whose squared distance is <= r^2 (r = the integer
 * part of min_peak_radius). out receives one tm_peak_info per peak (pixel_size converts the pixel
 * coordinates to x_pos / y_pos in Angstroms). No counterpart in the reference's GPU path (the peak
 * search runs on the host there). */
int tm_find_peaks(tm_core* h, long long total_ccfs, float threshold, int edge, float min_peak_radius, float pixel_size,
                  tm_peak_info* out, int capacity, int* n_peaks);

/* Allocation-free tm_find_peaks. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments and the same peaks, bit for bit, as tm_find_peaks; it calls no hipMalloc,
 * hipFree or allocating library routine. Its working memory is a static device array of
 * TM_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until the process ends (one bit
 * per search pixel for the exclusion discs, then the reduction partials), so the search image may
 * hold at most TM_NOALLOC_MAX_PIXELS pixels (4096 x 4096); a larger one gets TM_ERR_INVALID. Every
 * call on one GPU shares that GPU's array: calls that overlap in time on one GPU (from several
 * threads or handles) race, and the caller must serialise them. */
#define TM_NOALLOC_MAX_PIXELS (4096LL * 4096LL)
#define TM_NOALLOC_PARTS 1024
#define TM_NOALLOC_SCRATCH_BYTES (TM_NOALLOC_MAX_PIXELS / 8 + TM_NOALLOC_PARTS * 8 + 256)
int tm_find_peaks_noalloc(tm_core* h, long long total_ccfs, float threshold, int edge, float min_peak_radius,
                          float pixel_size, tm_peak_info* out, int capacity, int* n_peaks);

/* Checkpoint and resume of a search, with no counterpart in the reference (whose jobs are restarted
 * from their position range, match_template.cpp:378-379, 735-739).
 *
 * tm_run_partial runs orientations [first_orientation, first_orientation + n_orientations) of the
 * handle's range (indices in loop order, 0 .. number of orientations - 1) for one plane and folds
 * them into the maps. first_orientation 0 starts a new plane (size_i, defocus_i), as
 * tm_run_inner_loop does; a later first_orientation must equal the plane's next unsearched
 * orientation and the same (size_i, defocus_i). tm_run_inner_loop of a plane equals tm_run_partial
 * calls that cover it in order, bit for bit, however they split it.
 *
 * tm_state_size gives the bytes of the search state: MIP, best orientation index, FP64 sum and sum
 * of squares (W x H each), the histogram, the CCF count, the planes run so far and the next
 * orientation of the current plane. tm_export_state copies it to a caller buffer; tm_import_state
 * restores it into a handle initialised for the same search (same tm_init_params and inputs), so
 * that the search continues where the exporting handle stopped. */
int tm_run_partial(tm_core* h, const float* projection_filter, int size_i, int defocus_i, int first_orientation,
                   int n_orientations, long long* ccfs);
int tm_state_size(tm_core* h, long long* bytes);
int tm_export_state(tm_core* h, void* buffer, long long bytes);
int tm_import_state(tm_core* h, const void* buffer, long long bytes);

/* Graph mode, with no counterpart in the reference (its loop launches every kernel): when enabled,
 * the first tm_run_inner_loop after it captures the launches of one plane (all batches) into a
 * hipGraph and every later plane replays it; the plane's orientation-index offset is read from
 * device memory, written before each replay. Results equal the launch-by-launch path bit for bit.
 * Not used while timing is enabled. tm_init, and enabling or disabling, drop the captured graph. */
int tm_set_graph(tm_core* h, int enabled);

/* Groups: several handles on one device (typically one micrograph each, of any compiled sizes and
 * with their own position ranges) whose planes run as ONE hipGraph launch, with no counterpart in
 * the reference (which runs one image per process, src/gui/MatchTemplatePanel.cpp:674-778, and one
 * TemplateMatchingCore per thread, match_template.cpp:908-988). The handles are dealt to `lanes`
 * chains (largest work first, each to the chain with the least work so far): a chain runs its
 * handles one after another, and the chains are independent branches of the graph, so that `lanes`
 * searches share the GPU at a time (0: two). The graph is captured at the first tm_group_run and
 * replayed by the later ones, each plane's orientation-index offset read from device memory; a handle
 * re-initialised (tm_init) since the capture makes the next run capture again. Each handle's maps,
 * sums and histogram equal what tm_run_inner_loop with the same filter, size_i and defocus_i gives,
 * bit for bit, whatever the lanes.
 *
 * tm_group_create: handles[0..n) must be distinct and on one device, and must outlive the group.
 * tm_group_run: one plane for every handle: projection_filters[i], size_i[i], defocus_i[i] as in
 * tm_run_inner_loop; ccfs (may be NULL) receives each handle's CCF count. Handles with kernel timing
 * on, or with a plane left unfinished by tm_run_partial, give TM_ERR_STATE. */
typedef struct tm_group tm_group;
int tm_group_create(tm_core* const* handles, int n, int lanes, tm_group** out);
int tm_group_run(tm_group* g, const float* const* projection_filters, const int* size_i, const int* defocus_i,
                 long long* ccfs);
int tm_group_destroy(tm_group* g);

/* The launch shape tm_init chose: orientations per launch, the column pass in use (TM_COLUMNS_PAIRS
 * or TM_COLUMNS_PERSISTENT* at height 4096, TM_COLUMNS_AUTO elsewhere) and the column-tile width of
 * the inter-pass plane. Any pointer may be NULL. No counterpart in the reference. */
int tm_get_launch_shape(tm_core* h, int* batch_size, int* column_pass, int* z_tile);

/* Last error message of the calling thread. */
const char* tm_last_error(void);

/* Autotuner of the launch shape, with no counterpart in the reference: the batch depth
 * (orientations per launch) and the column-pass kernel (TM_COLUMNS_*: at height 4096 one block per
 * column pair with 4-column Z tiles, or the persistent pass with 2- or 4-column tiles; at other
 * heights column tiles of the compiled width or half of it). For each (column pass, batch)
 * candidate it builds a handle over a sample of the search (its first sample_orientations
 * orientations), runs one warm-up plane and one timed plane with projection_filter, and writes the
 * time per CCF in microseconds to us_per_ccf[column * n_batch + batch] (-1 where the column pass
 * does not apply to the search; us_per_ccf may be NULL). *best_batch and *best_column receive the
 * fastest pair (the earlier one on a tie). No candidate changes a result: each pixel's state is
 * updated in orientation order whatever the batch, and every column pass does the same arithmetic,
 * so all candidates give bit-identical maps, sums and histograms. */
int tm_autotune(int device_ordinal, const tm_init_params* params, const float* template_spectrum,
                const float* input_spectrum, const float* search_positions, const float* psi_list,
                const float* projection_filter, const int* batch_candidates, int n_batch, const int* column_candidates,
                int n_column, int sample_orientations, int* best_batch, int* best_column, double* us_per_ccf);

/* tm_autotune over batch candidates only, with params->column_pass as the one column pass. */
int tm_autotune_batch(int device_ordinal, const tm_init_params* params, const float* template_spectrum,
                      const float* input_spectrum, const float* search_positions, const float* psi_list,
                      const float* projection_filter, const int* candidates, int n_candidates, int sample_orientations,
                      int* best_batch, double* us_per_ccf);

#ifdef __cplusplus
}
#endif

#endif /* TM_HIP_H */
