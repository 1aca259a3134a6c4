/*
 * Caller-side preparation of the template-matching inputs (the host steps of the reference's
 * match_template that build the arguments of TemplateMatchingCore::Init and RunInnerLoop).
 * The 2-D and 3-D forward transforms run on the GPU; layouts as in tm_hip.h.
 * Every call returns TM_OK or a negative TM_ERR_* code (tm_prepare_last_error() explains).
 */
#ifndef TM_PREPARE_H
#define TM_PREPARE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Sizes after the reference's smooth-size factorisation and optional 90-degree rotation
 * (match_template.cpp:527-598): W x H of the searched image, whether it was rotated, the
 * number of whitening-curve points (match_template.cpp:699-700) and max_padding. */
int tm_prepare_sizes(int image_x, int image_y, int template_size, int* W, int* H, int* rotated, int* n_curve,
                     int* max_padding);

/* Input image -> whitened, normalised, quadrant-swapped spectrum (H x (W/2+1) complex) and the
 * whitening curve (x, y; n_curve points). match_template.cpp:521-598, 703-724. */
int tm_prepare_input_image(int device, const float* micrograph, int image_x, int image_y, int template_size,
                           float* spectrum, float* curve_x, float* curve_y);

/* Template volume (t^3, x fastest) -> padded (t * padding) spectrum, scaled 1/N, DC 0,
 * quadrant-swapped (match_template.cpp:621-625, 837-845). With spectrum == NULL only *P is set. */
int tm_prepare_template(int device, const float* volume, int t, float padding, float* spectrum, int* P);

/* The template spectrum of one pixel-size plane: as tm_prepare_template, with the padded volume
 * resampled by Image::ChangePixelSize(pixel_size_factor, tolerance 0.001) before the transform
 * (match_template.cpp:837-845, image.cpp:8705-8792); pixel_size_factor = (pixel_size + size_i *
 * pixel_size_step) / pixel_size. A factor within 0.001 of 1 gives tm_prepare_template's result. The
 * intermediate edge (1000 for P * factor < 1000) must have a compiled transform, else
 * TM_ERR_UNSUPPORTED; the new edge may be any even length up to 2048. */
int tm_prepare_template_scaled(int device, const float* volume, int t, float padding, float pixel_size_factor,
                               float* spectrum, int* P);

/* CTF x whitening projection filter, T x (T/2+1) complex (match_template.cpp:893-898). Defocus
 * values in Angstrom (defocus1/2 already include defocus_i * defocus_step). */
int tm_projection_filter(int T, const float* curve_x, const float* curve_y, int n_curve, float voltage_kV,
                         float spherical_aberration_mm, float amplitude_contrast, float defocus1, float defocus2,
                         float defocus_angle_deg, float pixel_size, float phase_shift_deg, float* out);

/* Euler search grid of EulerSearch::CalculateGridSearchPositions(false) with the C1 theta_max
 * override of match_template.cpp:689-697: {phi, theta} pairs; returns the count (writes at most
 * capacity pairs; out may be NULL to query). */
int tm_search_grid(char symmetry_type, int symmetry_number, float angular_step, float* out, int capacity);

/* In-plane angles of the inclusive float loop psi_start..psi_max (match_template.cpp:1001). */
int tm_psi_list(float psi_start, float psi_max, float psi_step, float* out, int capacity);

/* refine_template's input image: micrograph (image_x x image_y, x fastest, image_x even) -> outliers
 * replaced, whitened, normalised to unit spectral energy and transformed back to real space
 * (whitened, image_x x image_y dense), plus the whitening curve (*n_curve points into curve_x/curve_y).
 * refine_template.cpp:526-546. With whitened == NULL only *n_curve is set. The image sides must be
 * sizes with compiled transforms (TM_LINE_SIZES in tm_sizes.hpp), else TM_ERR_UNSUPPORTED. */
int tm_prepare_refine_image(int device, const float* micrograph, int image_x, int image_y, float* whitened,
                            float* curve_x, float* curve_y, int* n_curve);

const char* tm_prepare_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TM_PREPARE_H */
