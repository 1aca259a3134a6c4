// Side lengths with compiled kernels, and the launch shape of the per-size kernels.
//
// Image sides are what the reference's size policy produces (match_template.cpp:527-598 with
// ReturnClosestFactorizedUpper/Lower, functions.cpp:529-627): a power of two when one lies within
// 10% above, else the nearest 3-smooth (then 5-smooth) even size. Compiled here: powers of two from
// 256 to 4096; 1536, 2592, 2916, 3072 and 5832 (2048 x 1536, 2560^2 -> 2592^2, 4096 x 2880 -> 4096 x 2916,
// 3072 x 2048 rotated -> 2048 x 3072, a K3 frame 5760 x 4092 -> 5832 x 4096); 270, 324 and 384 (small
// 2/3/5-smooth sizes).
// Template and projection sides are powers of two.
#pragma once

#define TM_IMAGE_SIZES(X) \
  X(256) X(270) X(324) X(384) X(512) X(1024) X(1536) X(2048) X(2592) X(2916) X(3072) X(4096) X(5832)
#define TM_TEMPLATE_SIZES(X) X(32) X(64) X(128) X(256) X(512) X(1024)
// every line length a row or column transform of the input image, the template or the projection
// takes (the union of the two lists), and 1000, the intermediate edge of a template's pixel-size change
// (1 / tolerance 0.001)
#define TM_LINE_SIZES(X) \
  X(32) X(64) X(128) X(256) X(270) X(324) X(384) X(512) X(1000) X(1024) X(1536) X(2048) X(2592) X(2916) X(3072) \
  X(4096) X(5832)

namespace tmk {

// Column pass (k_mid, every height except 4096): columns per block and threads per column (the
// comments give each height's radix plan, tm_fft.hpp). Heights with radix-2/3/5/9 stages get fewer
// columns per block. The threads per column match the plan's butterfly counts (H / R per stage), so
// that no stage leaves half of them idle (3072 = 16 16 4 3 has 192 radix-16 butterflies: 192 threads,
// not 384) and more columns fit a CU. LDS per block: columns x (H + H/16) x 8 bytes.
constexpr int mid_cols(int H) {
  switch (H) {
    case 1536: return 4;
    case 2048: return 8;
    case 2592: return 2;
    case 2916: return 2;
    case 3072: return 2;
    case 5832: return 2;
    default: return 16;  // 256 .. 1024
  }
}
constexpr int mid_threads(int H) {
  switch (H) {
    case 256: return 16;    // 16 x 16
    case 270: return 64;    // 2 9 3 5
    case 324: return 64;    // 4 9 9
    case 384: return 32;    // 16 8 3
    case 512: return 32;    // 16 16 2
    case 1024: return 64;   // 16 16 4
    case 1536: return 96;   // 16 16 2 3
    case 2048: return 128;  // 16 16 8
    case 2592: return 324;  // 8 4 9 9 (mid_lr, tm_kernels.hpp)
    case 2916: return 324;  // 4 9 9 9
    case 3072: return 192;  // 16 16 4 3
    case 5832: return 486;  // 8 9 9 9
    default: return 0;
  }
}

// Row pass (k_last / k_last_dma): orientations per block for image width W. Each orientation has
// (W/2)/R1 threads, R1 the first radix of the W/2-point plan, and each thread owns R1/QR pixel pairs.
constexpr int last_qr(int W) {
  switch (W) {
    case 256: return 8;   // R1 8, 16 threads per orientation
    case 270: return 3;   // R1 3, 45
    case 324: return 2;   // R1 2, 81
    case 384: return 8;   // R1 8, 24
    case 512: return 8;   // R1 8, 32
    case 1024: return 4;  // R1 8, 64
    case 1536: return 4;  // R1 8, 96
    case 2048: return 4;  // R1 8, 128
    case 2592: return 2;  // R1 8, 162
    case 2916: return 1;  // R1 2, 729
    case 3072: return 2;  // R1 8, 192
    case 4096: return 2;  // R1 8, 256
    case 5832: return 1;  // R1 4, 729
    default: return 0;
  }
}

}  // namespace tmk
