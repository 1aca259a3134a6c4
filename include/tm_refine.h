/*
 * C ABI of the MI355X (gfx950) local refinement of template-matching peaks.
 *
 * It replaces the per-image work of the reference program refine_template
 * (src/programs/refine_template/refine_template.cpp:264-1229): the peak list taken from the scaled
 * MIP, one particle window per peak, TemplateScore (refine_template.cpp:25-80) and the coordinate
 * descent over defocus and the three Euler angles. TemplateScore calls of all peaks in flight run
 * together as one batch of device transforms; the descent logic of each peak runs on the host.
 *
 * TemplateScore resamples the projection with ChangePixelSize wherever the pixel-size factor differs
 * from 1 (the pixel-size scan and block, refine_template.cpp:931-1086, or a non-zero pixel-size map).
 * mask_radius > 0 masks each particle window as the program's CosineMask (refine_template.cpp:716).
 * The input image, template and maps are the program's inputs after its preparation
 * (tm_prepare_refine_image and tm_prepare_template in tm_prepare.h).
 *
 * Every call returns TM_OK (0) or a negative TM_ERR_* code from tm_hip.h; tm_refine_last_error()
 * describes the last failure on the calling thread. One handle per host thread.
 */
#ifndef TM_REFINE_H
#define TM_REFINE_H

#include "tm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tm_refine tm_refine;

typedef struct tm_refine_params {
  int image_x, image_y;         /* micrograph and map size (x even)                                   */
  int template_size;            /* t: windows and projections are t x t                                */
  int padded_size;              /* P: edge of the padded template spectrum (t * padding)               */
  float pixel_size;             /* Angstrom                                                            */
  float voltage_kV, spherical_aberration_mm, amplitude_contrast;
  float defocus1, defocus2, defocus_angle;  /* Angstrom, Angstrom, degrees                             */
  float phase_shift_deg;
  float angular_step, in_plane_angular_step;       /* descent steps (degrees)                          */
  float defocus_search_range, defocus_search_step; /* refine step = 2 * defocus_search_step            */
  float wanted_threshold;       /* peaks of the scaled MIP at or above this value are refined          */
  float min_peak_radius;        /* compared with the squared pixel distance, as the program does       */
  float xy_change_threshold;    /* Angstrom; output maps skip peaks that moved further when            */
  int exclude_above_xy_threshold; /* this flag is set                                                  */
  int shared_local_maps;        /* 1: the program's per-thread maps carry over from peak to peak, as in
                                   a run with one thread (peaks are then refined one at a time);
                                   0: every peak starts from the input maps, as a thread that handles
                                   one peak does (peaks are refined together)                          */
  int max_in_flight;            /* peaks refined together (the TemplateScore batch); 0 = 256           */
  float pixel_size_search_range, pixel_size_search_step; /* Angstrom; refine step = 2 * search step;
                                   TemplateScore resamples the projection by ChangePixelSize with
                                   factor (pixel_size + pixel size offset) / pixel_size               */
  float mask_radius;            /* Angstrom; > 0: each particle window gets the program's cosine mask
                                   (radius at most (t/2 - 1) pixels, 20 A falloff); 0: no mask         */
} tm_refine_params;

/* One refined peak (TemplateMatchFoundPeakInfo of all_peak_infos / all_peak_changes,
 * refine_template.cpp:1092-1113). Positions in pixels relative to the image centre. */
typedef struct tm_refined_peak {
  float x, y, value;                           /* found peak and its scaled-MIP value               */
  float psi, theta, phi, defocus, pixel_size;  /* the search's values at the found peak             */
  float x_pos, y_pos;                          /* refined position (found + shift)                  */
  float shift_x, shift_y;                      /* TemplateScore peak offset of the best score       */
  float r_psi, r_theta, r_phi, r_defocus, r_pixel_size; /* refined values (maps at best_address)    */
  float peak_height;                           /* scaled_mip * best / mip * t (best_scaled_mip)     */
  float mip;                                   /* best * t (best_mip)                               */
  float best_score;                            /* best TemplateScore value                          */
  float starting_score;                        /* score at the search's values, on the scaled scale */
  float offset;                                /* distance the peak moved (pixels)                  */
  long long best_address;                      /* padded-layout address ((x + 2) per row)            */
  int evaluations;                             /* TemplateScore calls for this peak                 */
  int reserved;
} tm_refined_peak;

/* Creates a handle bound to one device. */
int tm_refine_create(int device_ordinal, tm_refine** out);
int tm_refine_destroy(tm_refine* h);

/* Copies the refinement inputs to the device. whitened_image: image_x x image_y dense real (the
 * program's input_image after refine_template.cpp:526-546); template_spectrum: P^3 half-complex,
 * scaled 1/N, DC 0, quadrant-swapped (refine_template.cpp:497-507); curve_x/curve_y: the whitening
 * curve (n_curve points) that multiplies the CTF in each projection filter. */
int tm_refine_init(tm_refine* h, const tm_refine_params* params, const float* whitened_image,
                   const float* template_spectrum, const float* curve_x, const float* curve_y, int n_curve);

/* Replaces the whitened image and the whitening curve of an initialised handle (same image size
 * and curve length), keeping the template and the work buffers: the per-micrograph part of
 * tm_refine_init, for callers that score many micrographs of one size with one template. No
 * counterpart in the reference (its program handles one image per run). */
int tm_refine_set_image(tm_refine* h, const float* whitened_image, const float* curve_x, const float* curve_y,
                        int n_curve);

/* The per-micrograph preparation of refine_template.cpp:526-546 on the device, for a handle that
 * scores many micrographs of one size: the raw micrograph (image_x x image_y dense) is cleaned of
 * 5-sigma outliers, transformed, whitened by its own radial power spectrum and normalised, and
 * becomes the handle's whitened image and curve, as tm_prepare_refine_image followed by
 * tm_refine_set_image would give (the curve's bins are summed in FP64 rather than in float). No
 * counterpart in the reference (its program prepares one image per run, on the CPU). */
int tm_refine_set_micrograph(tm_refine* h, const float* micrograph);

/* Test hook: the handle's whitened image (dense) and whitening curve (n_curve points each); any
 * pointer may be NULL. */
int tm_refine_get_image(tm_refine* h, float* whitened, float* curve_x, float* curve_y);

/* Refines every peak of the scaled MIP. maps[7]: mip, scaled_mip, psi, theta, phi, defocus,
 * pixel_size, each image_x x image_y dense (the search's output maps). out_maps[7] (the array or
 * any entry may be NULL) receive best_mip, best_scaled_mip, best_psi, best_theta, best_phi,
 * best_defocus, best_pixel_size (dense) as the program writes them. peaks receives up to capacity
 * records in peak-list order; *n_peaks the number of peaks found. */
int tm_refine_run(tm_refine* h, const float* const* maps, float* const* out_maps, tm_refined_peak* peaks, int capacity,
                  int* n_peaks);

/* Test hooks with no counterpart in the reference: the peak list of a scaled MIP (dense;
 * out receives x, y, value per peak, at most capacity), and TemplateScore for the window at
 * (peak_x, peak_y) with the filter at defocus1|2 + current_defocus + step and the projection
 * resampled by pixel_size_factor (out: x, y, value). */
int tm_refine_peak_list(tm_refine* h, const float* scaled_mip, float* out, int capacity, int* n_peaks);
int tm_refine_score(tm_refine* h, int peak_x, int peak_y, float current_defocus, float step, float phi, float theta,
                    float psi, float pixel_size_factor, float* out);

/* Local exhaustive scoring, with no counterpart in the reference (the batched form of the
 * TemplateScore calls of a refinement around proposed locations): request i scores the window at
 * (peak_xy[2i], peak_xy[2i+1]) (pixels from the image centre) against the projection at
 * angles[3i..3i+2] (phi, theta, psi) with the filter at defocus1|2 + current_defocus and writes the
 * correlation peak's offset and value (x, y, value) to out[3i..3i+2], as tm_refine_score does.
 * Requests that share a window share one window transform; max_in_flight windows and requests
 * run per batch. */
int tm_refine_score_batch(tm_refine* h, int n, const int* peak_xy, const float* angles, float current_defocus,
                          float* out);

/* For t = 64 with P = t and no pixel-size change, a TemplateScore batch runs as one kernel with one
 * workgroup per score and every intermediate in LDS (on by default), and the batches of
 * tm_refine_score_batch overlap (the next batch's requests are prepared on the host while the device
 * runs the current one); results are bit-identical to the multi-kernel path, which enabled = 0
 * selects. No counterpart in the reference (a schedule). */
int tm_refine_set_fused(tm_refine* h, int enabled);

/* Per stage of the TemplateScore batches of the last tm_refine_run (extract_slice_cols,
 * slice_rows_c2r, projection_r2c, score_product, correlation_c2r, score_peak, score_fused), measured with HIP
 * events on the handle's stream when timing is enabled: total ms, number of launches and items
 * (TemplateScore calls) summed over the launches. Returns the number of entries written (at most cap).
 * No counterpart in the reference (diagnostics). */
int tm_refine_set_timing(tm_refine* h, int enabled);
int tm_refine_get_timing(tm_refine* h, const char** names, double* ms, long long* launches, long long* items, int cap);

/* Last error message of the calling thread. */
const char* tm_refine_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TM_REFINE_H */
