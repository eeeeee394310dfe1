// Fused passes of the Toeplitz-embedded SENSE normal operator (3-D, n x n
// planes with n = 64, 96 or 128): the first two (x and y zero-padded FFTs)
// and the last two (y and x inverse FFTs with crops, scale, output diagonals
// and the coil sum), each as one kernel.  Host interface.
#pragma once

#include "fft.hpp"

namespace hasty {

struct ToepCropSumArgs {
    const float2* grid;      // [B, nz, 2ny, 2nx], x positions class-major
    int64_t grid_b;          // elements per batch element
    float2* out;             // [nz, ny, nx]
    float2* partial;         // [groups, nz, ny, nx] scratch when groups > 1
    int64_t partial_elems;
    DiagOp<float2> m1, m2, io;   // element (b, row, x) at ptr[b*sb + row*sr + x], row = z*ny + y
    float scale;
    int nbatch, nz, groups, accumulate;
    int n;                   // plane size ny == nx
};

// First two passes fused: x and y zero-padded forward FFTs of
// in[b] (.) m1 (.) m2 into grid [B, nz, 2ny, 2nx] with x class-major.
struct ToepPadFftArgs {
    const float2* in;        // [B, nz, ny, nx]
    int64_t in_b;
    DiagOp<float2> m1, m2;   // element (b, row, x) at ptr[b*sb + row*sr + x]
    float2* grid;            // [B, nz, 2ny, 2nx], x positions class-major
    int64_t grid_b;
    int nbatch, nz;
    int n;                   // plane size ny == nx
};
void launch_toep_pad_fft(const ToepPadFftArgs& a, hipStream_t s);

// plane shapes the fused passes are built for (ny == nx, 64, 96 or 128)
bool toep_crop_sum_supported(int64_t ny, int64_t nx);
// groups: coil groups per plane (partial sums added by a second kernel)
void launch_toep_crop_sum(const ToepCropSumArgs& a, hipStream_t s);

}  // namespace hasty
