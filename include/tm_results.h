/*
 * Worker and master result records of a match_template run split over several jobs (the
 * reference's GUI job system): the float array a worker sends to the master
 * (match_template.cpp:1371-1531), the master's merge of those arrays
 * (AggregatedTemplateResult::AddResult, match_template.cpp:2015-2103) and its final scaling
 * (match_template.cpp:1617-1760). Host code over host arrays: byte layout and per-pixel
 * arithmetic, no device work. tm_merge (tm_hip.h) merges two device handles directly; these
 * records are for results that travel between processes or hosts.
 *
 * Record layout, in floats:
 *   [0] x, [1] y                  map size (the micrograph's original size)
 *   [2] n                         floats per map, (x + j) * y, j = 2 for even x, 1 for odd x
 *   [3] histogram bins            [4] CCFs computed by the worker   [5] sqrt(W * H) of the search image
 *   [6 + k n, 6 + (k + 1) n)      map k in the reference's padded real-space rows (x values, then j
 *                                 padding floats, written as 0): k = 0 MIP, 1 psi, 2 theta, 3 phi,
 *                                 4 defocus, 5 pixel size, 6 sum, 7 sum of squares
 *   [6 + 8 n, ...)                the histogram as int64 counts, two floats each (native byte order)
 *
 * Every call returns TM_OK (0) or a negative TM_ERR_* code from tm_hip.h; tm_result_last_error()
 * describes the last failure on the calling thread.
 */
#ifndef TM_RESULTS_H
#define TM_RESULTS_H

#include "tm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define TM_RESULT_META 6
#define TM_RESULT_MAPS 8

/* Floats in a record of an x by y map with histogram_bins bins (match_template.cpp:1376, 1452-1453). */
long long tm_result_size(int x, int y, int histogram_bins);

/* The worker's record (match_template.cpp:1371-1531). maps: 8 dense nx x ny row-major maps in the
 * micrograph's orientation (after the search's 90-degree rotation is undone), in the record's map
 * order; sums and sums of squares as floats. Each map is filled outside the valid area with its
 * average over a centred disc of radius 0.35 (nx + ny - 2 remove_npix_from_edge)
 * (Image::ReturnAverageOfRealValues, image.cpp:6263-6327): cut to (out_x - remove, out_y - remove)
 * (nx - remove / ny - remove where out_x > nx / out_y > ny), then padded to out_x x out_y with that
 * value, both centred (Image::Resize -> ClipInto, image.cpp:8522-8575, 8825-8838).
 * remove_npix_from_edge = round(particle radius / pixel size) (match_template.cpp:494). out must hold
 * tm_result_size(out_x, out_y, histogram_bins) floats (out_floats). */
int tm_result_pack(const float* const* maps, int nx, int ny, int out_x, int out_y, int remove_npix_from_edge,
                   const long long* histogram, int histogram_bins, float n_ccfs, double sqrt_input_pixels, float* out,
                   long long out_floats);

/* AggregatedTemplateResult::AddResult (match_template.cpp:2015-2103): folds a worker record into the
 * collated record. first != 0 starts the collated record (zeroed, header copied from result) and
 * sets *total_ccfs to 0 before adding. Per float of the maps: where result MIP > collated MIP
 * (strict, so equal values keep the earlier record), MIP, psi, theta, phi, defocus and pixel size
 * are taken from result; sums, sums of squares and histogram counts add; *total_ccfs += result[4]
 * in float. Records must agree in size (floats, n, bins). */
int tm_result_add(float* collated, const float* result, long long floats, float* total_ccfs, int first);

/* The master's scaling of a collated record (match_template.cpp:1617-1760), in float arithmetic as
 * there: mean = sum / total_ccfs, sd = sqrtf(sum_sq / total_ccfs - mean * mean); z = (MIP - mean) / sd
 * where sd > 0, else z = 0 and sd = 0. Outputs are dense x by y maps (any may be NULL):
 * mip_scaled = MIP * sqrt(W H), z, average = mean * sqrt(W H), std_dev = sd * sqrt(W H); angles
 * receives psi, theta, phi, defocus and pixel size (5 maps, in that order); histogram the counts. */
int tm_result_finalize(const float* collated, long long floats, float total_ccfs, float* mip_scaled, float* zscore,
                       float* average, float* std_dev, float* angles, long long* histogram);

/* Expected threshold of a search (the master's form, match_template.cpp:1799): sqrtf(2) * erfcinv(2 /
 * (x * y * total_ccfs)) with cisTEM's erfcinv (Winitzki's approximation of erfinv with a = 0.147,
 * functions.cpp:1282-1299), x and y the micrograph's original size; returned as the float the
 * reference computes. The local run discards the erfcinv result (match_template.cpp:1282); the
 * master's formula is used for both. */
double tm_expected_threshold(int image_x, int image_y, long long total_ccfs);

/* The histogram file's columns (match_template.cpp:1333-1364), bins 0..bins-1: snr[i] = histogram_min
 * + histogram_step / 2 + histogram_step * i (float), survival[i] = sum of counts of bins i..bins-1,
 * expected[i] = erfc(snr[i] / sqrtf(2)) / 2 * (sqrt_input_pixels^2 * total_ccfs), in float as there
 * (the reference's loop fills one entry past its array; here it stops at bins - 1). Any output may be
 * NULL. */
int tm_histogram_table(const long long* histogram, int bins, float histogram_min, float histogram_step,
                       double sqrt_input_pixels, long long total_ccfs, float* snr, double* survival,
                       double* expected);

const char* tm_result_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TM_RESULTS_H */
