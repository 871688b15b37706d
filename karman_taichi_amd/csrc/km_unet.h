// Internal interface of the learned p' preconditioner (km_unet.hip) used by
// the flexible BiCGSTAB (km_fbicg.hip).  Not part of the C-ABI.
#pragma once

#include "km_common.h"

namespace kmnet {

// 4-level U-Net, layers in the order e0 e1 e2 e3 d2 d1 d0 (3x3 convolutions)
// followed by a 1x1 output layer.  cin_p is the input channel count padded
// to a multiple of 8 (the MFMA operand loads move 8 BF16 channels at once).
constexpr int NL = 7;
struct LayerDef {
    int cin, cin_p, cout;
};
constexpr LayerDef kLayers[NL] = {{3, 8, 16},     {16, 16, 32}, {32, 32, 64}, {64, 64, 128},
                                  {192, 192, 64}, {96, 96, 32}, {48, 48, 16}};
constexpr int kOutIn = 16;

struct LayerOff {
    int cin, cin_p, cout, kp, ktp;   // kp = 9 cin_p and ktp = 9 cout, rounded up to 32
    int64_t w, b;                    // FP32 parameter offsets: W[cout][3][3][cin], b[cout]
    int64_t wf, wt;                  // BF16 offsets: forward [cout][kp], transposed-flipped [cin_p][ktp]
};
struct Layout {
    LayerOff L[NL];
    int64_t out_w, out_b, nparams, nwf, nwt;
};
Layout make_layout();

// Prepare the network for an NX x NY padded system and a batch of B samples
// (buffers grow only), and convert the FP32 parameters into the BF16 operand
// layouts on the context's stream.
int prepare(km_ctx* ctx, int B, int NX, int NY, const float* params);

// Per-solve statistics of one system (free-cell count and mean aP), from the
// FP32 coefficients.  Must precede apply() for that system.
int solve_stats(km_ctx* ctx, int NX, int NY, const float* coef, hipStream_t s);

// z = M(r) for one system (vectors in T = float or double); kernels return at once when
// *skip != 0 (skip may be null).  prepare() and solve_stats() must have run
// with the same parameters (biases and output weights are read from them).
template <typename T>
int apply(km_ctx* ctx, int NX, int NY, const float* coef, const T* r, T* z, const int* skip, hipStream_t s,
          const float* params);
// apply() with r.r given as G per-workgroup partial sums (written by the
// kernel that produced r), which saves the statistics pass over r.
// accumulate: z += M(r) on free cells (z keeps its other values) instead of z = M(r).
template <typename T>
int apply_pre(km_ctx* ctx, int NX, int NY, const float* coef, const T* r, T* z, const int* skip, hipStream_t s,
              const float* params, const double* rr_partials, int G, bool accumulate);

}  // namespace kmnet

// Frees the network buffers of a context (km_destroy).
void km_net_free(km_ctx* ctx);
