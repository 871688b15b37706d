// Learned p' preconditioner on gfx950: a 4-level U-Net of 3x3 convolutions
// with BF16 operands on the matrix cores (v_mfma_f32_16x16x32_bf16, FP32
// accumulation), its training backward pass, the residual loss and Adam.
//
// Tensors are channels-last, (B, H, W, C) with C contiguous, so one MFMA
// operand fragment (8 consecutive K values per lane) is one 16-byte load of
// 8 channels of one pixel.  A convolution is an implicit GEMM:
//   forward      out[co][px]  = sum_k Wf[co][k] X[px + tap(k)][c(k)]
//   input grad   dX[ci][px]   = sum_k Wt[ci][k] dY[px - tap][co(k)]   (flipped, transposed weights)
//   weight grad  dW[co][k]    = sum_px dY[px][co] X[px + tap(k)][c(k)]
// with k = tap * Cin + c.  The decoder's 2x nearest upsampling and channel
// concatenation are folded into the operand gather (no materialised
// concatenation); 2x2 average pools are small bandwidth-bound kernels.
#include "km_unet.h"

#include <cmath>

typedef __bf16 bf16;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

using kmnet::kLayers;
using kmnet::LayerOff;
using kmnet::Layout;
using kmnet::NL;

struct km_net {
    Layout lay{};
    LayerOff* dL = nullptr;     // device copy of lay.L
    int B = 0, NX = 0, NY = 0, H = 0, W = 0, Gs = 1;
    char* arena = nullptr;
    size_t bytes = 0;
    bf16 *wf = nullptr, *wt = nullptr;
    bf16 *X0, *e0, *p1, *e1, *p2, *e2, *p3, *e3, *d2, *d1, *d0;
    bf16 *dy_d0, *dy_d1, *dy_d2, *dy_e3, *dy_e2, *dy_e1, *dy_e0;
    float *z, *q, *ae, *dX0, *dX1, *dX2, *dp3, *dp2, *dp1, *wpart;
    float *c32, *r32, *z32;     // FP32 staging of one system (km_unet_apply)
    double *P, *stats, *loss;
    int64_t wpart_n = 0;
};

namespace {

constexpr int E0 = 0, E1 = 1, E2 = 2, E3 = 3, D2 = 4, D1 = 5, D0 = 6;
constexpr int kMaxBatch = 64;
constexpr int64_t kPartials = 1 << 20;          // doubles
constexpr int64_t kWpartMax = (int64_t)16 << 20;  // floats

__host__ __device__ constexpr int round32(int v) { return (v + 31) / 32 * 32; }

__device__ __forceinline__ bf16x8 zero8() {
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = (bf16)0.0f;
    return v;
}

__device__ __forceinline__ bool skipped(const int* skip) { return skip != nullptr && *skip != 0; }

__device__ __forceinline__ bool free_cell(const float* __restrict__ cb, int64_t n, int64_t k) {
    return cb[n + k] != 0.f || cb[2 * n + k] != 0.f || cb[3 * n + k] != 0.f || cb[4 * n + k] != 0.f;
}

// ------------------------------------------------------------ weights

// BF16 operand layouts from the FP32 parameters:
//   wf[co][k], k = tap * cin_p + c            (zero for c >= cin or k >= 9 cin_p)
//   wt[ci][k], k = tap' * cout + co, tap' = 8 - tap
__global__ void __launch_bounds__(KM_BLOCK)
k_prep_weights(const LayerOff* __restrict__ L, int64_t nwf, int64_t nwt, const float* __restrict__ params,
               bf16* __restrict__ wf, bf16* __restrict__ wt) {
    const int64_t n = nwf + nwt;
    for (int64_t idx = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * KM_BLOCK) {
        if (idx < nwf) {
            int l = 0;
            while (l + 1 < NL && idx >= L[l + 1].wf) ++l;
            const int64_t o = idx - L[l].wf;
            const int kp = L[l].kp, cin_p = L[l].cin_p, cin = L[l].cin;
            const int co = (int)(o / kp), k = (int)(o % kp);
            const int tap = k / cin_p, c = k % cin_p;
            float v = 0.f;
            if (tap < 9 && c < cin) v = params[L[l].w + ((int64_t)co * 9 + tap) * cin + c];
            wf[idx] = (bf16)v;
        } else {
            const int64_t j = idx - nwf;
            int l = 0;
            while (l + 1 < NL && j >= L[l + 1].wt) ++l;
            const int64_t o = j - L[l].wt;
            const int ktp = L[l].ktp, cout = L[l].cout, cin = L[l].cin;
            const int ci = (int)(o / ktp), k = (int)(o % ktp);
            const int tp = k / cout, co = k % cout;
            float v = 0.f;
            if (tp < 9 && ci < cin) v = params[L[l].w + ((int64_t)co * 9 + (8 - tp)) * cin + ci];
            wt[j] = (bf16)v;
        }
    }
}

// ------------------------------------------------------------ convolution

// One wave computes 16*Q pixels x 16*RTW output channels (row tiles
// blockIdx.z*RTW ...); a workgroup is 4 waves along the pixels, so the grid is
// (pixels / 64Q, batch, COUT / 16RTW) and low-resolution layers still launch
// thousands of waves.  The operand fragments of the next 32-wide K chunk are
// loaded while the MFMAs of the current one run (two register sets).
// MODE 0: input src0 (C0 channels) at the output resolution.
// MODE 1: input = [2x nearest upsample of src0 (C0 channels, half resolution) | src1 (C1 channels)].
// MODE 2: input = 2x2 average pool of src0 (C0 channels, double resolution), rounded to BF16 as k_pool does.
// OUTK 0: out = bf16(relu(acc + bias)); OUTK 1: out = acc in FP32 (input gradients);
// OUTK 2 (COUT 16, RTW 1): the network output fused in: the 1x1 layer over
// bf16(relu(acc + bias)) scaled as k_out does, written to the padded FP32 or
// FP64 field ``out`` (type T) on free interior cells (fo: output parameters,
// coef / stats / NX / NY: the system).
struct FusedOut {
    const float* pw;
    const float* coef;
    const double* stats;
    int NX, NY;
    int accum;    // add to the output field instead of overwriting it
};

template <int MODE, int C0, int C1, int COUT, int OUTK, int Q, int RTW, typename T = float>
__global__ void __launch_bounds__(KM_BLOCK)
k_conv3x3(int H, int W, const bf16* __restrict__ src0, const bf16* __restrict__ src1, const bf16* __restrict__ Wm,
          const float* __restrict__ bias, void* __restrict__ out, const int* __restrict__ skip, FusedOut fo) {
    if (skipped(skip)) return;
    constexpr int CIN = C0 + C1;
    constexpr int KP = round32(9 * CIN);
    constexpr int NCH = KP / 32;
    static_assert(C0 % 8 == 0 && C1 % 8 == 0 && COUT % (16 * RTW) == 0, "channel counts");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, col = lane & 15;
    const int64_t HW = (int64_t)H * W;
    const int b = blockIdx.y;
    const int rt0 = blockIdx.z * RTW;
    const int64_t px0 = ((int64_t)blockIdx.x * 4 + wave) * (16 * Q);
    if (px0 >= HW) return;
    int pi[Q], pj[Q];
    bool pv[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int64_t p = px0 + q * 16 + col;
        pv[q] = p < HW;
        pi[q] = pv[q] ? (int)(p / W) : 0;
        pj[q] = pv[q] ? (int)(p % W) : 0;
    }
    const bf16* s0 = src0 + (MODE == 0 ? (int64_t)b * HW * C0 : MODE == 1 ? (int64_t)b * (HW / 4) * C0
                                                                          : (int64_t)b * (4 * HW) * C0);
    const bf16* s1 = MODE == 1 ? src1 + (int64_t)b * HW * C1 : nullptr;
    const bf16* wrow = Wm + (int64_t)(rt0 * 16 + col) * KP + g * 8;
    const int Wc = W >> 1;
    auto load = [&](int kc, bf16x8 (&bv)[Q], bf16x8 (&av)[RTW]) {
        const int k0 = kc + g * 8;
        const int tap = k0 / CIN, c = k0 - tap * CIN;
        const int dy = tap / 3 - 1, dx = tap - (tap / 3) * 3 - 1;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int ii = pi[q] + dy, jj = pj[q] + dx;
            if (tap < 9 && pv[q] && ii >= 0 && ii < H && jj >= 0 && jj < W) {
                if (MODE == 2) {
                    const bf16* f = s0 + ((int64_t)(2 * ii) * (2 * W) + 2 * jj) * C0 + c;
                    const bf16x8 x00 = *reinterpret_cast<const bf16x8*>(f);
                    const bf16x8 x01 = *reinterpret_cast<const bf16x8*>(f + C0);
                    const bf16x8 x10 = *reinterpret_cast<const bf16x8*>(f + (int64_t)2 * W * C0);
                    const bf16x8 x11 = *reinterpret_cast<const bf16x8*>(f + (int64_t)2 * W * C0 + C0);
                    bf16x8 o;
#pragma unroll
                    for (int e = 0; e < 8; ++e)
                        o[e] = (bf16)(0.25f * ((((float)x00[e] + (float)x01[e]) + (float)x10[e]) + (float)x11[e]));
                    bv[q] = o;
                } else {
                    const bf16* ptr;
                    if (MODE == 0)
                        ptr = s0 + ((int64_t)ii * W + jj) * C0 + c;
                    else if (c < C0)
                        ptr = s0 + ((int64_t)(ii >> 1) * Wc + (jj >> 1)) * C0 + c;
                    else
                        ptr = s1 + ((int64_t)ii * W + jj) * C1 + (c - C0);
                    bv[q] = *reinterpret_cast<const bf16x8*>(ptr);
                }
            } else {
                bv[q] = zero8();
            }
        }
#pragma unroll
        for (int rt = 0; rt < RTW; ++rt) av[rt] = *reinterpret_cast<const bf16x8*>(wrow + (int64_t)rt * 16 * KP + kc);
    };
    floatx4 acc[RTW][Q];
#pragma unroll
    for (int rt = 0; rt < RTW; ++rt)
#pragma unroll
        for (int q = 0; q < Q; ++q) acc[rt][q] = floatx4{0.f, 0.f, 0.f, 0.f};
    auto mma = [&](const bf16x8 (&bv)[Q], const bf16x8 (&av)[RTW]) {
#pragma unroll
        for (int rt = 0; rt < RTW; ++rt)
#pragma unroll
            for (int q = 0; q < Q; ++q)
                acc[rt][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[rt], bv[q], acc[rt][q], 0, 0, 0);
    };
    bf16x8 b0[Q], a0[RTW], b1[Q], a1[RTW];
    load(0, b0, a0);
    for (int ch = 0; ch < NCH; ch += 2) {
        if (ch + 1 < NCH) load((ch + 1) * 32, b1, a1);
        mma(b0, a0);
        if (ch + 1 < NCH) {
            if (ch + 2 < NCH) load((ch + 2) * 32, b0, a0);
            mma(b1, a1);
        }
    }
    if constexpr (OUTK == 2) {
        static_assert(COUT == 16 && RTW == 1, "fused output needs all 16 channels in one wave");
        // lane (g, col) holds channels 4g .. 4g+3 of pixel col: gather the 16
        // activations into lane g = 0 and sum them in channel order
        const int64_t n = (int64_t)fo.NX * fo.NY;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            float y[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float a = acc[0][q][i] + bias[g * 4 + i];
                y[i] = (float)(bf16)(a > 0.f ? a : 0.f);
            }
            float all[16];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                all[i] = y[i];
                all[4 + i] = __shfl(y[i], col + 16, 64);
                all[8 + i] = __shfl(y[i], col + 32, 64);
                all[12 + i] = __shfl(y[i], col + 48, 64);
            }
            if (g == 0 && pv[q]) {
                const int64_t p = px0 + q * 16 + col;
                const int a = (int)(p / W), cc = (int)(p % W);
                if (a < fo.NX - 2 && cc < fo.NY - 2) {
                    const int64_t k = (int64_t)(a + 1) * fo.NY + (cc + 1);
                    if (free_cell(fo.coef, n, k)) {
                        float sum = 0.f;
#pragma unroll
                        for (int e = 0; e < 16; ++e) sum += (float)(bf16)fo.pw[e] * all[e];
                        sum += fo.pw[16];
                        const T v = (T)(float)(fo.stats[(int64_t)b * 8 + 4] * (double)sum);
                        T* o = static_cast<T*>(out) + (int64_t)b * n + k;
                        *o = fo.accum ? *o + v : v;
                    }
                }
            }
        }
        return;
    }
#pragma unroll
    for (int rt = 0; rt < RTW; ++rt) {
        const int co = (rt0 + rt) * 16 + g * 4;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            if (!pv[q]) continue;
            const int64_t o = ((int64_t)b * HW + px0 + q * 16 + col) * COUT + co;
            if (OUTK == 0) {
                bf16x4 v;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float y = acc[rt][q][i] + bias[co + i];
                    v[i] = (bf16)(y > 0.f ? y : 0.f);
                }
                *reinterpret_cast<bf16x4*>(static_cast<bf16*>(out) + o) = v;
            } else {
                *reinterpret_cast<floatx4*>(static_cast<float*>(out) + o) = acc[rt][q];
            }
        }
    }
}

// Weight gradient: one wave accumulates RT x CT output tiles (16 x 16) over
// its share of the pixels (K = pixels, 32 per MFMA); partial sums per wave go
// to part[(split * 4 + wave)][COUT][KP] and are summed in a fixed order by
// k_wgrad_reduce.  Workgroup (split, column group).
template <int MODE, int C0, int C1, int COUT, int CT>
__global__ void __launch_bounds__(KM_BLOCK)
k_wgrad(int H, int W, int B, const bf16* __restrict__ src0, const bf16* __restrict__ src1,
        const bf16* __restrict__ dY, int64_t pps, float* __restrict__ part) {
    constexpr int CIN = C0 + C1;
    constexpr int KP = round32(9 * CIN);
    constexpr int NCT = KP / 16;
    constexpr int RT = COUT / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, col = lane & 15;
    const int64_t HW = (int64_t)H * W, NPX = (int64_t)B * HW;
    const int split = blockIdx.x, cg = blockIdx.y;
    const int64_t wpps = pps / 4;
    const int64_t lo = (int64_t)split * pps + wave * wpps;
    const int64_t hi = lo + wpps < NPX ? lo + wpps : NPX;
    const int Wc = W >> 1;
    int cdy[CT], cdx[CT], cc[CT];
    bool cval[CT];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
        const int kcol = (cg * CT + ct) * 16 + col;
        const int tap = kcol / CIN;
        cc[ct] = kcol - tap * CIN;
        cval[ct] = (cg * CT + ct) < NCT && tap < 9;
        cdy[ct] = tap / 3 - 1;
        cdx[ct] = tap - (tap / 3) * 3 - 1;
    }
    floatx4 acc[RT][CT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) acc[rt][ct] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int64_t base = lo; base < hi; base += 32) {
        const int64_t P0 = base + g * 8;
        int b = 0, i = 0, j = 0;
        if (P0 < NPX) {
            b = (int)(P0 / HW);
            const int64_t rem = P0 - (int64_t)b * HW;
            i = (int)(rem / W);
            j = (int)(rem - (int64_t)i * W);
        }
        bf16x8 av[RT], bv[CT];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const bool ok = P0 + e < hi;
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) av[rt][e] = ok ? dY[(P0 + e) * COUT + rt * 16 + col] : (bf16)0.0f;
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) {
                bf16 val = (bf16)0.0f;
                const int ii = i + cdy[ct], jj = j + cdx[ct];
                if (ok && cval[ct] && ii >= 0 && ii < H && jj >= 0 && jj < W) {
                    const int c = cc[ct];
                    if (MODE == 0)
                        val = src0[(((int64_t)b * H + ii) * W + jj) * C0 + c];
                    else if (c < C0)
                        val = src0[((int64_t)b * (HW / 4) + (int64_t)(ii >> 1) * Wc + (jj >> 1)) * C0 + c];
                    else
                        val = src1[(((int64_t)b * H + ii) * W + jj) * C1 + (c - C0)];
                }
                bv[ct][e] = val;
            }
            if (++j == W) {
                j = 0;
                if (++i == H) { i = 0; ++b; }
            }
        }
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
                if (cg * CT + ct < NCT)
                    acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[rt], bv[ct], acc[rt][ct], 0, 0, 0);
    }
    float* dst = part + (int64_t)(split * 4 + wave) * COUT * KP;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
            if (cg * CT + ct >= NCT) continue;
            const int kcol = (cg * CT + ct) * 16 + col;
#pragma unroll
            for (int i = 0; i < 4; ++i) dst[(int64_t)(rt * 16 + g * 4 + i) * KP + kcol] = acc[rt][ct][i];
        }
}

__global__ void __launch_bounds__(KM_BLOCK)
k_wgrad_reduce(int nparts, int cout, int kp, int cin, int cin_p, const float* __restrict__ part, float* __restrict__ gw) {
    const int64_t n = (int64_t)cout * kp;
    for (int64_t idx = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * KM_BLOCK) {
        const int co = (int)(idx / kp), kcol = (int)(idx % kp);
        const int tap = kcol / cin_p, c = kcol % cin_p;
        if (tap >= 9 || c >= cin) continue;
        double s = 0.0;
        for (int p = 0; p < nparts; ++p) s += (double)part[(int64_t)p * n + idx];
        gw[((int64_t)co * 9 + tap) * cin + c] = (float)s;
    }
}

// Per-channel sums of a (npx, C) BF16 tensor: partials P[c * G + block].
template <int C>
__global__ void __launch_bounds__(KM_BLOCK) k_chan_sum(int64_t npx, const bf16* __restrict__ dY, double* __restrict__ P, int G) {
    __shared__ double sh[KM_BLOCK];
    constexpr int R = KM_BLOCK / C;
    const int c = threadIdx.x % C, r0 = threadIdx.x / C;
    const int64_t chunk = (npx + G - 1) / G;
    const int64_t lo = (int64_t)blockIdx.x * chunk;
    const int64_t hi = lo + chunk < npx ? lo + chunk : npx;
    double s = 0.0;
    for (int64_t p = lo + r0; p < hi; p += R) s += (double)(float)dY[p * C + c];
    sh[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < C) {
        double a = 0.0;
        for (int k = 0; k < R; ++k) a += sh[k * C + threadIdx.x];
        P[(int64_t)threadIdx.x * G + blockIdx.x] = a;
    }
}

// out[v] = sum_g P[v * G + g] (fixed order)
__global__ void __launch_bounds__(KM_BLOCK) k_reduce_rows(int nv, int G, const double* __restrict__ P, float* __restrict__ out) {
    const int v = blockIdx.x * KM_BLOCK + threadIdx.x;
    if (v >= nv) return;
    double s = 0.0;
    for (int g = 0; g < G; ++g) s += P[(int64_t)v * G + g];
    out[v] = (float)s;
}

// ------------------------------------------------------------ pools, splits

template <int C>
__global__ void __launch_bounds__(KM_BLOCK)
k_pool(int Ho, int Wo, int B, const bf16* __restrict__ in, bf16* __restrict__ out, const int* __restrict__ skip) {
    if (skipped(skip)) return;
    constexpr int NG = C / 8;
    const int64_t n = (int64_t)B * Ho * Wo * NG;
    const int Wi = 2 * Wo;
    for (int64_t idx = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * KM_BLOCK) {
        const int gc = (int)(idx % NG);
        int64_t p = idx / NG;
        const int j = (int)(p % Wo);
        p /= Wo;
        const int i = (int)(p % Ho);
        const int64_t b = p / Ho;
        const bf16* base = in + ((b * 2 * Ho + 2 * i) * (int64_t)Wi + 2 * j) * C + gc * 8;
        const bf16x8 x00 = *reinterpret_cast<const bf16x8*>(base);
        const bf16x8 x01 = *reinterpret_cast<const bf16x8*>(base + C);
        const bf16x8 x10 = *reinterpret_cast<const bf16x8*>(base + (int64_t)Wi * C);
        const bf16x8 x11 = *reinterpret_cast<const bf16x8*>(base + (int64_t)Wi * C + C);
        bf16x8 o;
#pragma unroll
        for (int e = 0; e < 8; ++e)
            o[e] = (bf16)(0.25f * ((((float)x00[e] + (float)x01[e]) + (float)x10[e]) + (float)x11[e]));
        *reinterpret_cast<bf16x8*>(out + ((b * Ho + i) * (int64_t)Wo + j) * C + gc * 8) = o;
    }
}

// Gradient of [up(coarse) | skip] -> gradient of the coarse activation:
// dy = bf16(sum over the 2x2 children of dX[..., c] if act > 0 else 0).
template <int CU, int CS>
__global__ void __launch_bounds__(KM_BLOCK)
k_split_up(int Hc, int Wc, int B, const float* __restrict__ dX, const bf16* __restrict__ act, bf16* __restrict__ dy) {
    constexpr int CT_ = CU + CS;
    const int64_t n = (int64_t)B * Hc * Wc * CU;
    const int Wf = 2 * Wc;
    for (int64_t idx = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * KM_BLOCK) {
        const int c = (int)(idx % CU);
        int64_t p = idx / CU;
        const int jc = (int)(p % Wc);
        p /= Wc;
        const int ic = (int)(p % Hc);
        const int64_t b = p / Hc;
        const float* d = dX + ((b * 2 * Hc + 2 * ic) * (int64_t)Wf + 2 * jc) * CT_ + c;
        const float s = ((d[0] + d[CT_]) + d[(int64_t)Wf * CT_]) + d[(int64_t)Wf * CT_ + CT_];
        dy[idx] = (bf16)((float)act[idx] > 0.f ? s : 0.f);
    }
}

// Encoder activation gradient: dy = bf16((skip gradient + unpool(dp)) if act > 0 else 0).
template <int C>
__global__ void __launch_bounds__(KM_BLOCK)
k_combine(int H, int W, int B, const float* __restrict__ skipg, int sstride, const float* __restrict__ dp,
          const bf16* __restrict__ act, bf16* __restrict__ dy) {
    const int64_t n = (int64_t)B * H * W * C;
    for (int64_t idx = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * KM_BLOCK) {
        const int c = (int)(idx % C);
        const int64_t p = idx / C;
        const int j = (int)(p % W);
        const int64_t pi = p / W;
        const int i = (int)(pi % H);
        const int64_t b = pi / H;
        const float sg = skipg[p * sstride + c];
        const float dv = dp[((b * (H / 2) + i / 2) * (int64_t)(W / 2) + j / 2) * C + c];
        const float v = sg + 0.25f * dv;
        dy[idx] = (bf16)((float)act[idx] > 0.f ? v : 0.f);
    }
}

// ------------------------------------------------------------ features, output

// Per system b: P[(b*3 + 0..2)*G + block] = partial sums of r.r and, when
// full, of aP and the count over free interior cells (a link is non-zero).
template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_net_stats(int NX, int NY, const T* __restrict__ r, int64_t rstride, const float* __restrict__ coef, int64_t cstride,
            int full, double* __restrict__ P, int G, const int* __restrict__ skip) {
    if (skipped(skip)) return;
    __shared__ double sh[4];
    const int b = blockIdx.y;
    const int64_t n = (int64_t)NX * NY;
    const int64_t chunk = (n + G - 1) / G;
    const int64_t lo = (int64_t)blockIdx.x * chunk;
    const int64_t hi = lo + chunk < n ? lo + chunk : n;
    const T* rb = r + b * rstride;
    const float* cb = coef + b * cstride;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int64_t k = lo + threadIdx.x; k < hi; k += KM_BLOCK) {
        const double v = (double)rb[k];
        s0 += v * v;
        if (full) {
            const int i = (int)(k / NY), j = (int)(k % NY);
            if (i >= 1 && i <= NX - 2 && j >= 1 && j <= NY - 2 && free_cell(cb, n, k)) {
                s1 += (double)cb[k];
                s2 += 1.0;
            }
        }
    }
    s0 = block_sum(s0, sh);
    s1 = block_sum(s1, sh);
    s2 = block_sum(s2, sh);
    if (threadIdx.x == 0) {
        P[((int64_t)b * 3 + 0) * G + blockIdx.x] = s0;
        P[((int64_t)b * 3 + 1) * G + blockIdx.x] = s1;
        P[((int64_t)b * 3 + 2) * G + blockIdx.x] = s2;
    }
}

// stats[b*8 + ...]: 0 r.r, 1 rs = sqrt(r.r / n_free), 3 abar, 4 rs / abar, 5 n_free.
__global__ void __launch_bounds__(KM_BLOCK)
k_net_stats_finish(int G, const double* __restrict__ P, int full, double* __restrict__ stats, const int* __restrict__ skip) {
    if (skipped(skip)) return;
    __shared__ double sh[4];
    const int b = blockIdx.x;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int k = threadIdx.x; k < G; k += KM_BLOCK) {
        a0 += P[((int64_t)b * 3 + 0) * G + k];
        if (full) {
            a1 += P[((int64_t)b * 3 + 1) * G + k];
            a2 += P[((int64_t)b * 3 + 2) * G + k];
        }
    }
    a0 = block_sum(a0, sh);
    a1 = block_sum(a1, sh);
    a2 = block_sum(a2, sh);
    if (threadIdx.x == 0) {
        double* S = stats + (int64_t)b * 8;
        if (full) {
            S[5] = a2;
            S[3] = a2 > 0.0 ? a1 / a2 : 1.0;
        }
        const double nf = S[5], abar = S[3];
        const double rs = nf > 0.0 ? sqrt(a0 / nf) : 0.0;
        S[0] = a0;
        S[1] = rs;
        S[4] = abar != 0.0 ? rs / abar : 0.0;
    }
}

// X0[b][a][c][0..7] = (r / rs, aP / abar on free cells, 1 on identity rows, 0 ...)
// zc (may be null): an output field of B x NX x NY values cleared on the way
// (the fused output layer then writes only free interior cells).
template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_features(int NX, int NY, int H, int W, int B, const T* __restrict__ r, int64_t rstride, const float* __restrict__ coef,
           int64_t cstride, const double* __restrict__ stats, bf16* __restrict__ X0, const int* __restrict__ skip,
           T* __restrict__ zc) {
    if (skipped(skip)) return;
    const int64_t HW = (int64_t)H * W, n = (int64_t)NX * NY, tot = (int64_t)B * HW;
    if (zc)
        for (int64_t i = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; i < (int64_t)B * n; i += (int64_t)gridDim.x * KM_BLOCK)
            zc[i] = T(0);
    for (int64_t idx = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * KM_BLOCK) {
        const int64_t b = idx / HW, p = idx % HW;
        const int a = (int)(p / W), c = (int)(p % W);
        float f0 = 0.f, f1 = 0.f, f2 = 0.f;
        if (a < NX - 2 && c < NY - 2) {
            const int64_t k = (int64_t)(a + 1) * NY + (c + 1);
            const float* cb = coef + b * cstride;
            const bool fr = free_cell(cb, n, k);
            const double rs = stats[b * 8 + 1], abar = stats[b * 8 + 3];
            if (rs > 0.0) f0 = (float)((double)r[b * rstride + k] / rs);
            f1 = fr ? (float)((double)cb[k] / abar) : 0.f;
            f2 = fr ? 0.f : 1.f;
        }
        bf16x8 o = zero8();
        o[0] = (bf16)f0;
        o[1] = (bf16)f1;
        o[2] = (bf16)f2;
        *reinterpret_cast<bf16x8*>(X0 + idx * 8) = o;
    }
}

// z = (rs / abar) (w_out . d0 + b_out) on free interior cells, 0 elsewhere
template <typename T>
__global__ void __launch_bounds__(KM_BLOCK)
k_out(int NX, int NY, int H, int W, int B, const bf16* __restrict__ d0, const float* __restrict__ pw,
      const float* __restrict__ coef, int64_t cstride, const double* __restrict__ stats, T* __restrict__ z,
      int64_t zstride, const int* __restrict__ skip) {
    if (skipped(skip)) return;
    const int64_t n = (int64_t)NX * NY, tot = (int64_t)B * n, HW = (int64_t)H * W;
    for (int64_t idx = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * KM_BLOCK) {
        const int64_t b = idx / n, k = idx % n;
        const int i = (int)(k / NY), j = (int)(k % NY);
        float v = 0.f;
        if (i >= 1 && i <= NX - 2 && j >= 1 && j <= NY - 2 && free_cell(coef + b * cstride, n, k)) {
            const bf16* dp = d0 + (b * HW + (int64_t)(i - 1) * W + (j - 1)) * 16;
            const bf16x8 lo = *reinterpret_cast<const bf16x8*>(dp);
            const bf16x8 hi = *reinterpret_cast<const bf16x8*>(dp + 8);
            float s = 0.f;
#pragma unroll
            for (int e = 0; e < 8; ++e) s += (float)(bf16)pw[e] * (float)lo[e];
#pragma unroll
            for (int e = 0; e < 8; ++e) s += (float)(bf16)pw[8 + e] * (float)hi[e];
            s += pw[16];
            v = (float)(stats[b * 8 + 4] * (double)s);
        }
        z[b * zstride + k] = (T)v;
    }
}

// ------------------------------------------------------------ loss

// Per system, FP32.  Residual loss (corr == null): q = r - A z, partials of
// q.q and r.r.  A-norm error loss: q = e = c - z, ae = A e, partials of e.ae
// and c.r.  P[(b * 2 + {0, 1}) * G + block].
__global__ void __launch_bounds__(KM_BLOCK)
k_loss_q(int NX, int NY, const float* __restrict__ coef, int64_t cstride, const float* __restrict__ r,
         const float* __restrict__ z, const float* __restrict__ corr, float* __restrict__ q, float* __restrict__ ae,
         double* __restrict__ P, int G) {
    __shared__ double sh[4];
    const int b = blockIdx.y;
    const int64_t n = (int64_t)NX * NY;
    const int64_t chunk = (n + G - 1) / G;
    const int64_t lo = (int64_t)blockIdx.x * chunk;
    const int64_t hi = lo + chunk < n ? lo + chunk : n;
    const float* cb = coef + b * cstride;
    const float* zb = z + b * n;
    const float* rb = r + b * n;
    float* qb = q + b * n;
    double s0 = 0.0, s1 = 0.0;
    if (corr == nullptr) {
        for (int64_t k = lo + threadIdx.x; k < hi; k += KM_BLOCK) {
            const int i = (int)(k / NY), j = (int)(k % NY);
            float az = cb[k] * zb[k];
            az = az + cb[n + k] * (i > 0 ? zb[k - NY] : 0.f);
            az = az + cb[2 * n + k] * (i < NX - 1 ? zb[k + NY] : 0.f);
            az = az + cb[3 * n + k] * (j < NY - 1 ? zb[k + 1] : 0.f);
            az = az + cb[4 * n + k] * (j > 0 ? zb[k - 1] : 0.f);
            const float qq = rb[k] - az;
            qb[k] = qq;
            s0 += (double)qq * (double)qq;
            s1 += (double)rb[k] * (double)rb[k];
        }
    } else {
        const float* ccb = corr + b * n;
        float* aeb = ae + b * n;
        for (int64_t k = lo + threadIdx.x; k < hi; k += KM_BLOCK) {
            const int i = (int)(k / NY), j = (int)(k % NY);
            const float e = ccb[k] - zb[k];
            float a = cb[k] * e;
            a = a + cb[n + k] * (i > 0 ? ccb[k - NY] - zb[k - NY] : 0.f);
            a = a + cb[2 * n + k] * (i < NX - 1 ? ccb[k + NY] - zb[k + NY] : 0.f);
            a = a + cb[3 * n + k] * (j < NY - 1 ? ccb[k + 1] - zb[k + 1] : 0.f);
            a = a + cb[4 * n + k] * (j > 0 ? ccb[k - 1] - zb[k - 1] : 0.f);
            qb[k] = e;
            aeb[k] = a;
            s0 += (double)e * (double)a;
            s1 += (double)ccb[k] * (double)rb[k];
        }
    }
    s0 = block_sum(s0, sh);
    s1 = block_sum(s1, sh);
    if (threadIdx.x == 0) {
        P[((int64_t)b * 2 + 0) * G + blockIdx.x] = s0;
        P[((int64_t)b * 2 + 1) * G + blockIdx.x] = s1;
    }
}

// loss[b] = num / den; the denominator goes to stats[b*8 + 6] for the gradient.
__global__ void __launch_bounds__(KM_BLOCK)
k_loss_finish(int G, const double* __restrict__ P, double* __restrict__ stats, double* __restrict__ loss) {
    __shared__ double sh[4];
    const int b = blockIdx.x;
    double a = 0.0, d = 0.0;
    for (int k = threadIdx.x; k < G; k += KM_BLOCK) {
        a += P[((int64_t)b * 2 + 0) * G + k];
        d += P[((int64_t)b * 2 + 1) * G + k];
    }
    a = block_sum(a, sh);
    d = block_sum(d, sh);
    if (threadIdx.x == 0) {
        stats[(int64_t)b * 8 + 6] = d;
        loss[b] = d > 0.0 ? a / d : 0.0;
    }
}

// dL/dnet on free interior cells, (rs/abar) times
//   residual loss:  -2 A^T q / (den B)
//   A-norm loss:    -(A e + A^T e) / (den B)     (q holds e, ae holds A e)
// fused with the backward of the 1x1 output layer: dy_d0 = bf16(dnet w_out
// [d0 > 0]) and partial sums of dnet * d0[c] (16) and dnet (1) into P[v * G + block].
__global__ void __launch_bounds__(KM_BLOCK)
k_loss_grad(int NX, int NY, int H, int W, int B, const float* __restrict__ coef, int64_t cstride,
            const float* __restrict__ q, const float* __restrict__ ae, const double* __restrict__ stats,
            const bf16* __restrict__ d0, const float* __restrict__ pw, bf16* __restrict__ dy, double* __restrict__ P,
            int G) {
    __shared__ double sh[17][4];
    const int64_t n = (int64_t)NX * NY, HW = (int64_t)H * W, tot = (int64_t)B * HW;
    float wb[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) wb[c] = (float)(bf16)pw[c];
    double acc[17];
#pragma unroll
    for (int v = 0; v < 17; ++v) acc[v] = 0.0;
    for (int64_t idx = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; idx < tot; idx += (int64_t)G * KM_BLOCK) {
        const int64_t b = idx / HW, p = idx % HW;
        const int a = (int)(p / W), c = (int)(p % W);
        float dnet = 0.f;
        if (a < NX - 2 && c < NY - 2) {
            const int64_t k = (int64_t)(a + 1) * NY + (c + 1);
            const float* cb = coef + b * cstride;
            if (free_cell(cb, n, k)) {
                const float* qb = q + b * n;
                float at = cb[k] * qb[k];
                at = at + cb[n + k + NY] * qb[k + NY];
                at = at + cb[2 * n + k - NY] * qb[k - NY];
                at = at + cb[3 * n + k - 1] * qb[k - 1];
                at = at + cb[4 * n + k + 1] * qb[k + 1];
                const double den = stats[b * 8 + 6], sc = stats[b * 8 + 4];
                if (ae == nullptr)
                    dnet = (float)(sc * (-2.0 / (den * (double)B)) * (double)at);
                else
                    dnet = (float)(sc * (-1.0 / (den * (double)B)) * ((double)ae[b * n + k] + (double)at));
            }
        }
        const bf16x8 lo = *reinterpret_cast<const bf16x8*>(d0 + idx * 16);
        const bf16x8 hi = *reinterpret_cast<const bf16x8*>(d0 + idx * 16 + 8);
        bf16x8 olo, ohi;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float x0 = (float)lo[e], x1 = (float)hi[e];
            olo[e] = (bf16)(x0 > 0.f ? dnet * wb[e] : 0.f);
            ohi[e] = (bf16)(x1 > 0.f ? dnet * wb[8 + e] : 0.f);
            acc[e] += (double)dnet * (double)x0;
            acc[8 + e] += (double)dnet * (double)x1;
        }
        acc[16] += (double)dnet;
        *reinterpret_cast<bf16x8*>(dy + idx * 16) = olo;
        *reinterpret_cast<bf16x8*>(dy + idx * 16 + 8) = ohi;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int v = 0; v < 17; ++v) {
        const double s = wave_sum(acc[v]);
        if (lane == 0) sh[v][w] = s;
    }
    __syncthreads();
    if (threadIdx.x < 17) {
        const int v = threadIdx.x;
        P[(int64_t)v * G + blockIdx.x] = (sh[v][0] + sh[v][1]) + (sh[v][2] + sh[v][3]);
    }
}

__global__ void __launch_bounds__(KM_BLOCK)
k_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
       float lr, float b1, float b2, float eps, float bc1, float bc2) {
    for (int64_t i = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * KM_BLOCK) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * (gi * gi);
        m[i] = mi;
        v[i] = vi;
        const float mh = mi / bc1, vh = vi / bc2;
        p[i] = p[i] - lr * mh / (sqrtf(vh) + eps);
    }
}

template <typename S, typename D>
__global__ void __launch_bounds__(KM_BLOCK) k_convert(int64_t n, const S* __restrict__ a, D* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * KM_BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * KM_BLOCK)
        out[i] = (D)a[i];
}

// ------------------------------------------------------------ host side

int flat_blocks(int64_t n) {
    int64_t g = (n + KM_BLOCK - 1) / KM_BLOCK;
    if (g > 4096) g = 4096;
    return g < 1 ? 1 : (int)g;
}

template <int MODE, int C0, int C1, int COUT, int OUTK, int Q, int RTW, typename T = float>
void conv(int H, int W, int B, const bf16* s0, const bf16* s1, const bf16* Wm, const float* bias, void* out,
          const int* skip, hipStream_t st, FusedOut fo = FusedOut{}) {
    const int64_t HW = (int64_t)H * W;
    const dim3 grid((unsigned)((HW + 64 * Q - 1) / (64 * Q)), (unsigned)B, (unsigned)(COUT / (16 * RTW)));
    hipLaunchKernelGGL((k_conv3x3<MODE, C0, C1, COUT, OUTK, Q, RTW, T>), grid, dim3(KM_BLOCK), 0, st, H, W, s0, s1,
                       Wm, bias, out, skip, fo);
}

template <int C>
void pool(int Ho, int Wo, int B, const bf16* in, bf16* out, const int* skip, hipStream_t st) {
    hipLaunchKernelGGL(k_pool<C>, dim3(flat_blocks((int64_t)B * Ho * Wo * (C / 8))), dim3(KM_BLOCK), 0, st, Ho, Wo, B,
                       in, out, skip);
}

// Weight gradient of layer l: partials, then the fixed-order reduction into
// the FP32 gradient vector, plus the bias gradient.
template <int MODE, int C0, int C1, int COUT>
void wgrad(km_net* N, int l, int H, int W, int B, const bf16* s0, const bf16* s1, const bf16* dY, float* grad,
           hipStream_t st) {
    constexpr int CIN = C0 + C1, KP = round32(9 * CIN), NCT = KP / 16;
    constexpr int CT = COUT >= 128 ? 4 : 8;
    const int ncg = (NCT + CT - 1) / CT;
    const int64_t npx = (int64_t)B * H * W;
    int64_t nsplit = (1024 + ncg - 1) / ncg;
    const int64_t cap = kWpartMax / (4LL * COUT * KP);
    if (nsplit > cap) nsplit = cap;
    const int64_t most = (npx + 127) / 128;
    if (nsplit > most) nsplit = most;
    if (nsplit < 1) nsplit = 1;
    int64_t pps = (npx + nsplit - 1) / nsplit;
    pps = (pps + 127) / 128 * 128;
    nsplit = (npx + pps - 1) / pps;
    hipLaunchKernelGGL((k_wgrad<MODE, C0, C1, COUT, CT>), dim3((unsigned)nsplit, (unsigned)ncg), dim3(KM_BLOCK), 0, st,
                       H, W, B, s0, s1, dY, pps, N->wpart);
    const LayerOff& L = N->lay.L[l];
    hipLaunchKernelGGL(k_wgrad_reduce, dim3(flat_blocks((int64_t)COUT * KP)), dim3(KM_BLOCK), 0, st,
                       (int)(nsplit * 4), COUT, KP, L.cin, L.cin_p, (const float*)N->wpart, grad + L.w);
    const int G = (int)(npx / 4096 + 1 < 1024 ? npx / 4096 + 1 : 1024);
    hipLaunchKernelGGL(k_chan_sum<COUT>, dim3(G), dim3(KM_BLOCK), 0, st, npx, dY, N->P, G);
    hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(KM_BLOCK), 0, st, COUT, G, (const double*)N->P, grad + L.b);
}

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

int alloc(km_ctx* ctx, km_net* N, int B, int NX, int NY) {
    if (N->arena && B <= N->B && NX == N->NX && NY == N->NY) return KM_OK;
    KM_REQUIRE(ctx, B >= 1 && B <= kMaxBatch, "batch %d outside [1, %d]", B, kMaxBatch);
    KM_REQUIRE(ctx, NX >= 4 && NY >= 4, "system %d x %d too small for the network", NX, NY);
    if (N->arena && NX == N->NX && NY == N->NY && B < N->B) B = N->B;
    const int H = (NX - 2 + 7) / 8 * 8, W = (NY - 2 + 7) / 8 * 8;
    const int64_t hw0 = (int64_t)H * W, hw1 = hw0 / 4, hw2 = hw0 / 16, hw3 = hw0 / 64;
    const int64_t n = (int64_t)NX * NY;
    const size_t bb = sizeof(bf16), fb = sizeof(float);
    struct Item { void** ptr; size_t bytes; };
    km_net& M = *N;
    Item items[] = {
        {(void**)&M.X0, B * hw0 * 8 * bb},   {(void**)&M.e0, B * hw0 * 16 * bb},  {(void**)&M.p1, B * hw1 * 16 * bb},
        {(void**)&M.e1, B * hw1 * 32 * bb},  {(void**)&M.p2, B * hw2 * 32 * bb},  {(void**)&M.e2, B * hw2 * 64 * bb},
        {(void**)&M.p3, B * hw3 * 64 * bb},  {(void**)&M.e3, B * hw3 * 128 * bb}, {(void**)&M.d2, B * hw2 * 64 * bb},
        {(void**)&M.d1, B * hw1 * 32 * bb},  {(void**)&M.d0, B * hw0 * 16 * bb},
        {(void**)&M.dy_d0, B * hw0 * 16 * bb}, {(void**)&M.dy_d1, B * hw1 * 32 * bb},
        {(void**)&M.dy_d2, B * hw2 * 64 * bb}, {(void**)&M.dy_e3, B * hw3 * 128 * bb},
        {(void**)&M.dy_e2, B * hw2 * 64 * bb}, {(void**)&M.dy_e1, B * hw1 * 32 * bb},
        {(void**)&M.dy_e0, B * hw0 * 16 * bb},
        {(void**)&M.z, B * n * fb},          {(void**)&M.q, B * n * fb},          {(void**)&M.ae, B * n * fb},
        {(void**)&M.dX0, B * hw0 * 48 * fb}, {(void**)&M.dX1, B * hw1 * 96 * fb}, {(void**)&M.dX2, B * hw2 * 192 * fb},
        {(void**)&M.dp3, B * hw3 * 64 * fb}, {(void**)&M.dp2, B * hw2 * 32 * fb}, {(void**)&M.dp1, B * hw1 * 16 * fb},
        {(void**)&M.wpart, kWpartMax * fb},
        {(void**)&M.c32, 5 * n * fb},        {(void**)&M.r32, n * fb},            {(void**)&M.z32, n * fb},
        {(void**)&M.P, kPartials * sizeof(double)}, {(void**)&M.stats, kMaxBatch * 8 * sizeof(double)},
        {(void**)&M.loss, kMaxBatch * sizeof(double)},
        {(void**)&M.wf, (size_t)M.lay.nwf * bb}, {(void**)&M.wt, (size_t)M.lay.nwt * bb},
        {(void**)&M.dL, sizeof(LayerOff) * NL},
    };
    size_t total = 0;
    for (const auto& it : items) total += align256(it.bytes);
    KM_HIP(ctx, hipSetDevice(ctx->device));
    if (N->arena) {
        KM_HIP_SYNC(ctx, hipStreamSynchronize(ctx->stream));
        km_drop_graphs(ctx);     // captured graphs hold pointers into the arena
        KM_HIP_SYNC(ctx, hipFree(N->arena));
        N->arena = nullptr;
        N->bytes = 0;
    }
    void* p = nullptr;
    hipError_t e;
    {
        KmHostLock lk(km_host_mutex());
        e = hipMalloc(&p, total);
    }
    if (e != hipSuccess) return km_fail(ctx, KM_ENOMEM, "network hipMalloc(%zu) failed: %s", total, hipGetErrorString(e));
    N->arena = static_cast<char*>(p);
    N->bytes = total;
    size_t off = 0;
    for (const auto& it : items) {
        *it.ptr = N->arena + off;
        off += align256(it.bytes);
    }
    N->B = B;
    N->NX = NX;
    N->NY = NY;
    N->H = H;
    N->W = W;
    int64_t gs = n / 4096 + 1;
    N->Gs = (int)(gs > 256 ? 256 : gs);
    KM_HIP_SYNC(ctx, hipMemcpy(N->dL, N->lay.L, sizeof(LayerOff) * NL, hipMemcpyHostToDevice));
    KM_HIP_SYNC(ctx, hipMemset(N->stats, 0, kMaxBatch * 8 * sizeof(double)));
    return KM_OK;
}

km_net* get_net(km_ctx* ctx) {
    if (!ctx->net) {
        ctx->net = new km_net();
        ctx->net->lay = kmnet::make_layout();
    }
    return ctx->net;
}

void forward(km_net* N, int B, const float* params, const int* skip, hipStream_t st) {
    const int H = N->H, W = N->W;
    const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4, H3 = H / 8, W3 = W / 8;
    const LayerOff* L = N->lay.L;
    const bf16* wf = N->wf;
    conv<0, 8, 0, 16, 0, 4, 1>(H, W, B, N->X0, nullptr, wf + L[E0].wf, params + L[E0].b, N->e0, skip, st);
    pool<16>(H1, W1, B, N->e0, N->p1, skip, st);
    conv<0, 16, 0, 32, 0, 4, 1>(H1, W1, B, N->p1, nullptr, wf + L[E1].wf, params + L[E1].b, N->e1, skip, st);
    pool<32>(H2, W2, B, N->e1, N->p2, skip, st);
    conv<0, 32, 0, 64, 0, 1, 2>(H2, W2, B, N->p2, nullptr, wf + L[E2].wf, params + L[E2].b, N->e2, skip, st);
    pool<64>(H3, W3, B, N->e2, N->p3, skip, st);
    conv<0, 64, 0, 128, 0, 1, 2>(H3, W3, B, N->p3, nullptr, wf + L[E3].wf, params + L[E3].b, N->e3, skip, st);
    conv<1, 128, 64, 64, 0, 1, 1>(H2, W2, B, N->e3, N->e2, wf + L[D2].wf, params + L[D2].b, N->d2, skip, st);
    conv<1, 64, 32, 32, 0, 2, 1>(H1, W1, B, N->d2, N->e1, wf + L[D1].wf, params + L[D1].b, N->d1, skip, st);
    conv<1, 32, 16, 16, 0, 4, 1>(H, W, B, N->d1, N->e0, wf + L[D0].wf, params + L[D0].b, N->d0, skip, st);
}

// Inference: pools folded into the next convolution's gather and the output
// layer into d0's epilogue (same arithmetic as forward() + k_out).
template <typename T>
void forward_infer(km_net* N, const float* params, const int* skip, hipStream_t st, T* z, const float* coef, int NX,
                   int NY, int accum = 0) {
    const int H = N->H, W = N->W;
    const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4, H3 = H / 8, W3 = W / 8;
    const LayerOff* L = N->lay.L;
    const bf16* wf = N->wf;
    conv<0, 8, 0, 16, 0, 4, 1>(H, W, 1, N->X0, nullptr, wf + L[E0].wf, params + L[E0].b, N->e0, skip, st);
    conv<2, 16, 0, 32, 0, 4, 1>(H1, W1, 1, N->e0, nullptr, wf + L[E1].wf, params + L[E1].b, N->e1, skip, st);
    conv<2, 32, 0, 64, 0, 1, 2>(H2, W2, 1, N->e1, nullptr, wf + L[E2].wf, params + L[E2].b, N->e2, skip, st);
    conv<2, 64, 0, 128, 0, 1, 2>(H3, W3, 1, N->e2, nullptr, wf + L[E3].wf, params + L[E3].b, N->e3, skip, st);
    conv<1, 128, 64, 64, 0, 1, 1>(H2, W2, 1, N->e3, N->e2, wf + L[D2].wf, params + L[D2].b, N->d2, skip, st);
    conv<1, 64, 32, 32, 0, 2, 1>(H1, W1, 1, N->d2, N->e1, wf + L[D1].wf, params + L[D1].b, N->d1, skip, st);
    const FusedOut fo{params + N->lay.out_w, coef, (const double*)N->stats, NX, NY, accum};
    conv<1, 32, 16, 16, 2, 4, 1, T>(H, W, 1, N->d1, N->e0, wf + L[D0].wf, params + L[D0].b, z, skip, st, fo);
}

// Backward from dy_d0 (already written by k_loss_grad) to every parameter.
void backward(km_net* N, int B, float* grad, hipStream_t st) {
    const int H = N->H, W = N->W;
    const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4, H3 = H / 8, W3 = W / 8;
    const LayerOff* L = N->lay.L;
    const bf16* wt = N->wt;
    // d0: input [up(d1) | e0]
    wgrad<1, 32, 16, 16>(N, D0, H, W, B, N->d1, N->e0, N->dy_d0, grad, st);
    conv<0, 16, 0, 48, 1, 4, 1>(H, W, B, N->dy_d0, nullptr, wt + L[D0].wt, nullptr, N->dX0, nullptr, st);
    hipLaunchKernelGGL((k_split_up<32, 16>), dim3(flat_blocks((int64_t)B * H1 * W1 * 32)), dim3(KM_BLOCK), 0, st, H1, W1,
                       B, (const float*)N->dX0, (const bf16*)N->d1, N->dy_d1);
    // d1: input [up(d2) | e1]
    wgrad<1, 64, 32, 32>(N, D1, H1, W1, B, N->d2, N->e1, N->dy_d1, grad, st);
    conv<0, 32, 0, 96, 1, 2, 1>(H1, W1, B, N->dy_d1, nullptr, wt + L[D1].wt, nullptr, N->dX1, nullptr, st);
    hipLaunchKernelGGL((k_split_up<64, 32>), dim3(flat_blocks((int64_t)B * H2 * W2 * 64)), dim3(KM_BLOCK), 0, st, H2, W2,
                       B, (const float*)N->dX1, (const bf16*)N->d2, N->dy_d2);
    // d2: input [up(e3) | e2]
    wgrad<1, 128, 64, 64>(N, D2, H2, W2, B, N->e3, N->e2, N->dy_d2, grad, st);
    conv<0, 64, 0, 192, 1, 1, 2>(H2, W2, B, N->dy_d2, nullptr, wt + L[D2].wt, nullptr, N->dX2, nullptr, st);
    hipLaunchKernelGGL((k_split_up<128, 64>), dim3(flat_blocks((int64_t)B * H3 * W3 * 128)), dim3(KM_BLOCK), 0, st, H3,
                       W3, B, (const float*)N->dX2, (const bf16*)N->e3, N->dy_e3);
    // e3: input p3
    wgrad<0, 64, 0, 128>(N, E3, H3, W3, B, N->p3, nullptr, N->dy_e3, grad, st);
    conv<0, 128, 0, 64, 1, 1, 1>(H3, W3, B, N->dy_e3, nullptr, wt + L[E3].wt, nullptr, N->dp3, nullptr, st);
    hipLaunchKernelGGL(k_combine<64>, dim3(flat_blocks((int64_t)B * H2 * W2 * 64)), dim3(KM_BLOCK), 0, st, H2, W2, B,
                       (const float*)(N->dX2 + 128), 192, (const float*)N->dp3, (const bf16*)N->e2, N->dy_e2);
    // e2: input p2
    wgrad<0, 32, 0, 64>(N, E2, H2, W2, B, N->p2, nullptr, N->dy_e2, grad, st);
    conv<0, 64, 0, 32, 1, 1, 1>(H2, W2, B, N->dy_e2, nullptr, wt + L[E2].wt, nullptr, N->dp2, nullptr, st);
    hipLaunchKernelGGL(k_combine<32>, dim3(flat_blocks((int64_t)B * H1 * W1 * 32)), dim3(KM_BLOCK), 0, st, H1, W1, B,
                       (const float*)(N->dX1 + 64), 96, (const float*)N->dp2, (const bf16*)N->e1, N->dy_e1);
    // e1: input p1
    wgrad<0, 16, 0, 32>(N, E1, H1, W1, B, N->p1, nullptr, N->dy_e1, grad, st);
    conv<0, 32, 0, 16, 1, 4, 1>(H1, W1, B, N->dy_e1, nullptr, wt + L[E1].wt, nullptr, N->dp1, nullptr, st);
    hipLaunchKernelGGL(k_combine<16>, dim3(flat_blocks((int64_t)B * H * W * 16)), dim3(KM_BLOCK), 0, st, H, W, B,
                       (const float*)(N->dX0 + 32), 48, (const float*)N->dp1, (const bf16*)N->e0, N->dy_e0);
    // e0: input X0
    wgrad<0, 8, 0, 16>(N, E0, H, W, B, N->X0, nullptr, N->dy_e0, grad, st);
}

}  // namespace

namespace kmnet {

Layout make_layout() {
    Layout lay{};
    int64_t off = 0, wf = 0, wt = 0;
    for (int l = 0; l < NL; ++l) {
        LayerOff& L = lay.L[l];
        L.cin = kLayers[l].cin;
        L.cin_p = kLayers[l].cin_p;
        L.cout = kLayers[l].cout;
        L.kp = round32(9 * L.cin_p);
        L.ktp = round32(9 * L.cout);
        L.w = off;
        off += (int64_t)L.cout * 9 * L.cin;
        L.b = off;
        off += L.cout;
        L.wf = wf;
        wf += (int64_t)L.cout * L.kp;
        L.wt = wt;
        wt += (int64_t)L.cin_p * L.ktp;
    }
    lay.out_w = off;
    off += kOutIn;
    lay.out_b = off;
    off += 1;
    lay.nparams = off;
    lay.nwf = wf;
    lay.nwt = wt;
    return lay;
}

int prepare(km_ctx* ctx, int B, int NX, int NY, const float* params) {
    km_net* N = get_net(ctx);
    if (int rc = alloc(ctx, N, B, NX, NY)) return rc;
    hipLaunchKernelGGL(k_prep_weights, dim3(flat_blocks(N->lay.nwf + N->lay.nwt)), dim3(KM_BLOCK), 0, ctx->stream,
                       (const LayerOff*)N->dL, N->lay.nwf, N->lay.nwt, params, N->wf, N->wt);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

int solve_stats(km_ctx* ctx, int NX, int NY, const float* coef, hipStream_t s) {
    km_net* N = ctx->net;
    KM_REQUIRE(ctx, N && N->arena && N->NX == NX && N->NY == NY, "network not prepared for %d x %d", NX, NY);
    const int64_t n = (int64_t)NX * NY;
    hipLaunchKernelGGL(k_net_stats<float>, dim3(N->Gs, 1), dim3(KM_BLOCK), 0, s, NX, NY, coef, (int64_t)0, coef,
                       (int64_t)0, 1, N->P, N->Gs, (const int*)nullptr);
    hipLaunchKernelGGL(k_net_stats_finish, dim3(1), dim3(KM_BLOCK), 0, s, N->Gs, (const double*)N->P, 1, N->stats,
                       (const int*)nullptr);
    (void)n;
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

template <typename T>
int apply(km_ctx* ctx, int NX, int NY, const float* coef, const T* r, T* z, const int* skip, hipStream_t s,
          const float* params) {
    km_net* N = ctx->net;
    KM_REQUIRE(ctx, N && N->arena && N->NX == NX && N->NY == NY, "network not prepared for %d x %d", NX, NY);
    hipLaunchKernelGGL(k_net_stats<T>, dim3(N->Gs, 1), dim3(KM_BLOCK), 0, s, NX, NY, r, (int64_t)0, coef,
                       (int64_t)0, 0, N->P, N->Gs, skip);
    hipLaunchKernelGGL(k_net_stats_finish, dim3(1), dim3(KM_BLOCK), 0, s, N->Gs, (const double*)N->P, 0, N->stats, skip);
    hipLaunchKernelGGL(k_features<T>, dim3(flat_blocks((int64_t)N->H * N->W)), dim3(KM_BLOCK), 0, s, NX, NY, N->H,
                       N->W, 1, r, (int64_t)0, coef, (int64_t)0, (const double*)N->stats, N->X0, skip, z);
    forward_infer<T>(N, params, skip, s, z, coef, NX, NY);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

// The same with r.r already reduced to G partials by the kernel that wrote r.
template <typename T>
int apply_pre(km_ctx* ctx, int NX, int NY, const float* coef, const T* r, T* z, const int* skip, hipStream_t s,
              const float* params, const double* rr_partials, int G, bool accumulate) {
    km_net* N = ctx->net;
    KM_REQUIRE(ctx, N && N->arena && N->NX == NX && N->NY == NY, "network not prepared for %d x %d", NX, NY);
    hipLaunchKernelGGL(k_net_stats_finish, dim3(1), dim3(KM_BLOCK), 0, s, G, rr_partials, 0, N->stats, skip);
    hipLaunchKernelGGL(k_features<T>, dim3(flat_blocks((int64_t)N->H * N->W)), dim3(KM_BLOCK), 0, s, NX, NY, N->H,
                       N->W, 1, r, (int64_t)0, coef, (int64_t)0, (const double*)N->stats, N->X0, skip,
                       accumulate ? (T*)nullptr : z);
    forward_infer<T>(N, params, skip, s, z, coef, NX, NY, accumulate ? 1 : 0);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

template int apply_pre<float>(km_ctx*, int, int, const float*, const float*, float*, const int*, hipStream_t,
                              const float*, const double*, int, bool);
template int apply_pre<double>(km_ctx*, int, int, const float*, const double*, double*, const int*, hipStream_t,
                               const float*, const double*, int, bool);
template int apply<float>(km_ctx*, int, int, const float*, const float*, float*, const int*, hipStream_t, const float*);
template int apply<double>(km_ctx*, int, int, const float*, const double*, double*, const int*, hipStream_t,
                           const float*);

}  // namespace kmnet

void km_net_free(km_ctx* ctx) {
    if (!ctx || !ctx->net) return;
    KmHostLock lk(km_host_mutex());
    if (ctx->net->arena) (void)hipFree(ctx->net->arena);
    delete ctx->net;
    ctx->net = nullptr;
}

extern "C" {

int km_unet_param_count(int64_t* n) {
    if (!n) return KM_EINVAL;
    *n = kmnet::make_layout().nparams;
    return KM_OK;
}

int km_unet_apply(km_ctx* ctx, const float* params, int NX, int NY, const double* coef, const double* r, double* z) {
    if (!ctx) return KM_EINVAL;
    KM_REQUIRE(ctx, params && coef && r && z, "null pointer");
    if (int rc = kmnet::prepare(ctx, 1, NX, NY, params)) return rc;
    km_net* N = ctx->net;
    const int64_t n = (int64_t)NX * NY;
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL((k_convert<double, float>), dim3(flat_blocks(5 * n)), dim3(KM_BLOCK), 0, s, 5 * n, coef, N->c32);
    hipLaunchKernelGGL((k_convert<double, float>), dim3(flat_blocks(n)), dim3(KM_BLOCK), 0, s, n, r, N->r32);
    KM_CHECK_LAUNCH(ctx);
    if (int rc = kmnet::solve_stats(ctx, NX, NY, N->c32, s)) return rc;
    if (int rc = kmnet::apply<float>(ctx, NX, NY, N->c32, N->r32, N->z32, nullptr, s, params)) return rc;
    hipLaunchKernelGGL((k_convert<float, double>), dim3(flat_blocks(n)), dim3(KM_BLOCK), 0, s, n, (const float*)N->z32, z);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

int km_unet_loss_grad(km_ctx* ctx, const float* params, int B, int NX, int NY, const float* coef, const float* r,
                      const float* corr, float* grad, double* loss) {
    if (!ctx) return KM_EINVAL;
    KM_REQUIRE(ctx, params && coef && r && grad, "null pointer");
    if (int rc = kmnet::prepare(ctx, B, NX, NY, params)) return rc;
    km_net* N = ctx->net;
    const int64_t n = (int64_t)NX * NY;
    const int H = N->H, W = N->W, Gs = N->Gs;
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(k_net_stats<float>, dim3(Gs, B), dim3(KM_BLOCK), 0, s, NX, NY, r, n, coef, 5 * n, 1, N->P, Gs,
                       (const int*)nullptr);
    hipLaunchKernelGGL(k_net_stats_finish, dim3(B), dim3(KM_BLOCK), 0, s, Gs, (const double*)N->P, 1, N->stats,
                       (const int*)nullptr);
    hipLaunchKernelGGL(k_features<float>, dim3(flat_blocks((int64_t)B * H * W)), dim3(KM_BLOCK), 0, s, NX, NY, H, W, B,
                       r, n, coef, 5 * n, (const double*)N->stats, N->X0, (const int*)nullptr, (float*)nullptr);
    forward(N, B, params, nullptr, s);
    hipLaunchKernelGGL(k_out<float>, dim3(flat_blocks((int64_t)B * n)), dim3(KM_BLOCK), 0, s, NX, NY, H, W, B,
                       (const bf16*)N->d0, params + N->lay.out_w, coef, 5 * n, (const double*)N->stats, N->z, n,
                       (const int*)nullptr);
    hipLaunchKernelGGL(k_loss_q, dim3(Gs, B), dim3(KM_BLOCK), 0, s, NX, NY, coef, 5 * n, r, (const float*)N->z, corr,
                       N->q, N->ae, N->P, Gs);
    hipLaunchKernelGGL(k_loss_finish, dim3(B), dim3(KM_BLOCK), 0, s, Gs, (const double*)N->P, N->stats, N->loss);
    int64_t gl = (int64_t)B * H * W / 4096 + 1;
    const int Gl = (int)(gl > 1024 ? 1024 : gl);
    hipLaunchKernelGGL(k_loss_grad, dim3(Gl), dim3(KM_BLOCK), 0, s, NX, NY, H, W, B, coef, 5 * n, (const float*)N->q,
                       corr ? (const float*)N->ae : nullptr, (const double*)N->stats, (const bf16*)N->d0,
                       params + N->lay.out_w, N->dy_d0, N->P, Gl);
    hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(KM_BLOCK), 0, s, 17, Gl, (const double*)N->P, grad + N->lay.out_w);
    backward(N, B, grad, s);
    KM_CHECK_LAUNCH(ctx);
    if (loss) {
        KM_HIP(ctx, hipMemcpyAsync(ctx->h_res + 32, N->loss, B * sizeof(double), hipMemcpyDeviceToHost, s));
        KM_HIP_SYNC(ctx, hipStreamSynchronize(s));
        double m = 0.0;
        for (int b = 0; b < B; ++b) m += ctx->h_res[32 + b];
        *loss = m / B;
    }
    return KM_OK;
}

int km_adam_step(km_ctx* ctx, int64_t n, float* params, const float* grad, float* m, float* v, int step, double lr,
                 double b1, double b2, double eps) {
    if (!ctx) return KM_EINVAL;
    KM_REQUIRE(ctx, params && grad && m && v && n > 0 && step >= 1, "invalid Adam arguments");
    const double bc1 = 1.0 - pow(b1, step), bc2 = 1.0 - pow(b2, step);
    hipLaunchKernelGGL(k_adam, dim3(flat_blocks(n)), dim3(KM_BLOCK), 0, ctx->stream, n, params, grad, m, v, (float)lr,
                       (float)b1, (float)b2, (float)eps, (float)bc1, (float)bc2);
    KM_CHECK_LAUNCH(ctx);
    return KM_OK;
}

}  // extern "C"
