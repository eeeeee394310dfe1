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
    // frames (blockIdx.y): frame f reads grid + f*grid_f, writes out + f*out_f
    // and partial + f*partial_f, and its diagonals at ptr + f*sf
    int nframes = 1;
    int64_t grid_f = 0, out_f = 0, partial_f = 0;
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
    // frames (blockIdx.y): frame f reads in + f*in_f and its diagonals at
    // ptr + f*sf, and writes grid + f*grid_f
    int nframes = 1;
    int64_t in_f = 0, grid_f = 0;
};
void launch_toep_pad_fft(const ToepPadFftArgs& a, hipStream_t s);

// NUFFT type-2 fine grid: the x and y zero-padded forward passes of a 3-D c64
// plan with n_y = n_x = 128 on 256 x 256 (centred modes, shift -64), fused
// (nufft_fused.hip); img [B, nz, 128, 128] -> grid [B, nz, 256, 256]
struct NufftFwdXyArgs {
    const float2* in;
    int64_t in_b;
    DiagOp<float2> m1, m2;                         // load diagonals, as the row pass takes them
    const float* dz; const float* dy; const float* dx;   // deapodisation (all or none)
    float2* grid;
    int64_t grid_b;
    int nbatch, nz, sign, groups;
};
bool nufft_fwd_xy_supported(int64_t ny, int64_t nx, int64_t Ly, int64_t Lx, int shift_y, int shift_x);
void launch_nufft_fwd_xy(const NufftFwdXyArgs& a, hipStream_t s);

// NUFFT type-1 fine grid: the y and x inverse passes with their crops of the
// same plans, fused (nufft_fused.hip); grid [B, nz, 256, 256] (after the z
// pass) -> img [B, nz, 128, 128] with the row pass's epilogue: scale, the
// store diagonals, deapodisation, accumulate (no batch sum)
struct NufftInvXyArgs {
    const float2* in;
    int64_t in_b;
    float2* out;
    int64_t out_b;
    float scale;
    DiagOp<float2> m1, m2, io;                     // store diagonals, as the row pass takes them
    const float* dz; const float* dy; const float* dx;   // deapodisation (all or none)
    int accumulate;
    int nbatch, nz, sign, groups;
};
bool nufft_inv_xy_supported(int64_t ny, int64_t nx, int64_t Ly, int64_t Lx, int shift_y, int shift_x);
void launch_nufft_inv_xy(const NufftInvXyArgs& a, hipStream_t s);

// plane shapes the fused passes are built for (ny == nx, 64, 96 or 128)
bool toep_crop_sum_supported(int64_t ny, int64_t nx);
// groups: coil groups per plane (partial sums added by a second kernel)
void launch_toep_crop_sum(const ToepCropSumArgs& a, hipStream_t s);

}  // namespace hasty
