// Batched in-place 2D complex FFT on the in-LDS engine (see fft2d.hip).
#pragma once

#include "common.h"

namespace nusht {

// `batch` row-major n0 x n1 arrays, forward (sign -1) or backward (sign +1, unnormalised).
void fft2d_c2c(double2* data, int n0, int n1, int batch, int backward, hipStream_t st);
void fft2d_c2c(float2* data, int n0, int n1, int batch, int backward, hipStream_t st);

}  // namespace nusht
