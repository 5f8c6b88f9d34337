// Learned upper bound for the DPBF solve: a two-layer mean-aggregation message-passing network
// (hidden 64, BF16 weights and activations, FP32 accumulation) scores vertices; the DPBF optimum
// of the subgraph induced by the top-scored vertices (plus every group member) is a feasible
// tree, whose cost bounds the full solve.  The network never decides the answer: the bounded
// full DPBF returns the exact optimum, and without a feasible subgraph tree the plain DPBF runs.
//
//   x_v   = [log1p(deg)/8, log1p(min incident weight)/8, member of group 0..7, 0 x 6]   (16, bf16)
//   h1_v  = relu(W1 [x_v ; mean_{u in N(v)} x_u] + b1)                                    (64)
//   h2_v  = relu(W2 [h1_v ; mean_{u in N(v)} h1_u] + b2)                                   (64)
//   s_v   = w3 . h2_v + b3
// Dense layers run on the gfx950 MFMA unit (v_mfma_f32_16x16x32_bf16); aggregation is an
// HBM-bound gather.  Training: weighted logistic loss on "vertex is in the optimal tree" labels
// produced by the GPU DPBF itself, gradients by hand-written backward kernels, Adam in FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <condition_variable>
#include <deque>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

#include "gst_batch.h"

// A subgraph holding more than this many per mille of the graph's arcs is not solved: its DPBF
// would cost as much as the bounded full solve it is meant to speed up (on R-MAT graphs the top-scored
// vertices are the hubs, and their induced subgraph holds most of the arcs).  Two such subgraphs on
// one graph and the model stops selecting on it: its queries run with the spider bound alone.
#ifndef GST_SUB_ARC_PERMILLE
#define GST_SUB_ARC_PERMILLE 250
#endif
#ifndef GST_SELECT_LANES
#define GST_SELECT_LANES 2   // selection streams of the batched path, each with its own activations
#endif
#ifndef GST_SELECT_DEPTH
#define GST_SELECT_DEPTH 4   // selections queued at once over all selection streams
#endif
// forward dense layers: workgroups per CU (each wave holds all of W in registers, loaded once, so
// fewer waves with more row tiles each read W less often)
#ifndef GST_DENSE_GRID
#define GST_DENSE_GRID 2
#endif
// selected queries held ready for the engine (16: +0.8% on the BA V=200k learned stream over 8)
#ifndef GST_SELECT_AHEAD
#define GST_SELECT_AHEAD 16
#endif
#include "gst_internal.h"
#include "gst_mpnn.h"

#define HIPCHK(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return gst_set_error(ctx, GST_HIP_ERROR, "%s failed: %s (%s:%d)", #call,         \
                                 hipGetErrorString(e_), __FILE__, __LINE__);                 \
    } while (0)

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int FIN = 16;    // input features (10 used)
constexpr int HID = 64;
constexpr int K1 = 2 * FIN;   // 32
constexpr int K2 = 2 * HID;   // 128
constexpr int AGG_HUB = 512;  // rows longer than this are aggregated in chunks

// parameter offsets (floats)
constexpr int OFF_W1 = 0;                    // [HID][K1]
constexpr int OFF_B1 = OFF_W1 + HID * K1;    // [HID]
constexpr int OFF_W2 = OFF_B1 + HID;         // [HID][K2]
constexpr int OFF_B2 = OFF_W2 + HID * K2;    // [HID]
constexpr int OFF_W3 = OFF_B2 + HID;         // [HID]
constexpr int OFF_B3 = OFF_W3 + HID;         // [1]
constexpr int NPARAM = OFF_B3 + 1;

__device__ __forceinline__ float bf2f(bf16 x) { return (float)x; }
__device__ __forceinline__ bf16 f2bf(float x) { return (bf16)x; }  // round to nearest even

// ---------------------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------------------
// Rows of an uploaded graph run in increasing weight, so the lightest incident arc is the first.
// Two 16-byte stores per vertex.
__global__ void k_features(const uint32_t* row, const uint32_t* w, int w_float, const uint32_t* gm, uint32_t n,
                           bf16* x) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const uint32_t rs = row[v], re = row[v + 1];
        float minw = 0.f;
        if (re > rs) minw = w_float ? __uint_as_float(w[rs]) : (float)w[rs];
        const uint32_t gmv = gm[v];
        bf16x8 a, b;
        a[0] = f2bf(log1pf((float)(re - rs)) * 0.125f);
        a[1] = f2bf(log1pf(minw) * 0.125f);
#pragma unroll
        for (int i = 0; i < 6; ++i) a[2 + i] = f2bf(((gmv >> i) & 1u) ? 1.f : 0.f);
        b[0] = f2bf(((gmv >> 6) & 1u) ? 1.f : 0.f);
        b[1] = f2bf(((gmv >> 7) & 1u) ? 1.f : 0.f);
#pragma unroll
        for (int i = 2; i < 8; ++i) b[i] = f2bf(0.f);
        bf16x8* xv = (bf16x8*)(x + (uint64_t)v * FIN);
        xv[0] = a;
        xv[1] = b;
    }
}

// ---------------------------------------------------------------------------------------
// aggregation (HBM-bound gathers)
//   forward : out[u] = mean_{nb in N(u)} in[nb]             (bf16 in/out, fp32 sums)
//   backward: out[u] += sum_{nb in N(u)} in[nb] / deg(nb)   (fp32; transpose of the mean, the
//             graph being symmetric)
// One wave per light row: lanes = (neighbour slot, feature pair).  Rows longer than AGG_HUB are
// split into AGG_HUB-arc chunks whose partial sums are added atomically into fp32 scratch.
// ---------------------------------------------------------------------------------------
// forward: lanes = (neighbour slot, 8-feature chunk); each lane gathers 16 bytes (8 bf16) of a
// neighbour row per step, two steps in flight, and the slots are summed by shuffles
template <int D>
__global__ void __launch_bounds__(256) k_aggregate_fwd(const uint32_t* row, const int32_t* col, uint32_t n,
                                                       const bf16* in, bf16* out, float* hub_acc,
                                                       const uint32_t* hub_rows, uint32_t n_hub_items) {
    constexpr int CH = D / 8;           // 16-byte chunks per row
    constexpr int SLOTS = 64 / CH;      // neighbours per step
    const uint32_t lane = __lane_id();
    const uint32_t ch = lane % CH, slot = lane / CH;
    const uint32_t gwave = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const uint32_t nwaves = gridDim.x * blockDim.x / 64;
    const uint32_t total = n + n_hub_items;
    for (uint32_t it = gwave; it < total; it += nwaves) {
        uint32_t u, e0, e1, h = 0;
        const bool hub = it >= n;
        if (!hub) {
            u = it;
            e0 = row[u];
            e1 = row[u + 1];
            if (e1 - e0 > AGG_HUB) continue;
        } else {
            const uint32_t i = it - n;
            u = hub_rows[2 * i];
            h = hub_rows[2 * i + 1] & 0xFFFFFu;
            const uint32_t c = hub_rows[2 * i + 1] >> 20;
            const uint32_t rs = row[u], re = row[u + 1];
            e0 = rs + c * AGG_HUB;
            e1 = min(re, e0 + AGG_HUB);
        }
        float acc[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[j] = 0.f;
        for (uint32_t e = e0 + slot; e < e1; e += 2 * SLOTS) {
            const uint32_t nb0 = (uint32_t)col[e];
            const bool two = e + SLOTS < e1;
            const uint32_t nb1 = two ? (uint32_t)col[e + SLOTS] : 0u;
            const bf16x8 x0 = *(const bf16x8*)(in + (uint64_t)nb0 * D + 8 * ch);
            bf16x8 x1;
            if (two) x1 = *(const bf16x8*)(in + (uint64_t)nb1 * D + 8 * ch);
#pragma unroll
            for (int j = 0; j < 8; ++j) acc[j] += bf2f(x0[j]) + (two ? bf2f(x1[j]) : 0.f);
        }
#pragma unroll
        for (int o = CH; o < 64; o <<= 1) {
#pragma unroll
            for (int j = 0; j < 8; ++j) acc[j] += __shfl_xor(acc[j], o, 64);
        }
        if (slot == 0) {
            if (!hub) {
                const uint32_t deg = e1 - e0;
                const float inv = deg ? 1.f / (float)deg : 0.f;
                bf16x8 o;
#pragma unroll
                for (int j = 0; j < 8; ++j) o[j] = f2bf(acc[j] * inv);
                *(bf16x8*)(out + (uint64_t)u * D + 8 * ch) = o;
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) atomicAdd(&hub_acc[(uint64_t)h * D + 8 * ch + j], acc[j]);
            }
        }
    }
}

template <int D, bool BWD>
__global__ void __launch_bounds__(256) k_aggregate_bwd(const uint32_t* row, const int32_t* col, uint32_t n, const void* in_,
                                                   void* out_, float* hub_acc, const uint32_t* hub_rows,
                                                   uint32_t n_hub_items) {
    constexpr int FP = D / 2;            // feature pairs
    constexpr int SLOTS = 64 / FP;       // neighbour slots per wave
    const uint32_t lane = __lane_id();
    const uint32_t fp = lane % FP, slot = lane / FP;
    const uint32_t gwave = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const uint32_t nwaves = gridDim.x * blockDim.x / 64;
    // light rows
    for (uint32_t u = gwave; u < n; u += nwaves) {
        const uint32_t rs = row[u], re = row[u + 1];
        const uint32_t deg = re - rs;
        if (deg > AGG_HUB) continue;
        float s0 = 0.f, s1 = 0.f;
        for (uint32_t e = rs + slot; e < re; e += SLOTS) {
            const uint32_t nb = (uint32_t)col[e];
            if (BWD) {
                const float* in = (const float*)in_;
                const float sc = 1.f / (float)(row[nb + 1] - row[nb]);
                s0 += in[(uint64_t)nb * D + 2 * fp] * sc;
                s1 += in[(uint64_t)nb * D + 2 * fp + 1] * sc;
            } else {
                const bf16* in = (const bf16*)in_;
                s0 += bf2f(in[(uint64_t)nb * D + 2 * fp]);
                s1 += bf2f(in[(uint64_t)nb * D + 2 * fp + 1]);
            }
        }
#pragma unroll
        for (int o = FP; o < 64; o <<= 1) {
            s0 += __shfl_xor(s0, o, 64);
            s1 += __shfl_xor(s1, o, 64);
        }
        if (slot == 0) {
            if (BWD) {
                float* out = (float*)out_;
                out[(uint64_t)u * D + 2 * fp] += s0;
                out[(uint64_t)u * D + 2 * fp + 1] += s1;
            } else {
                bf16* out = (bf16*)out_;
                const float inv = deg ? 1.f / (float)deg : 0.f;
                out[(uint64_t)u * D + 2 * fp] = f2bf(s0 * inv);
                out[(uint64_t)u * D + 2 * fp + 1] = f2bf(s1 * inv);
            }
        }
    }
    // hub chunks: item i -> (hub index h, chunk c) packed as h | c << 32 in hub_rows pairs
    for (uint32_t i = gwave; i < n_hub_items; i += nwaves) {
        const uint32_t u = hub_rows[2 * i];
        const uint32_t h = hub_rows[2 * i + 1] & 0xFFFFFu;
        const uint32_t c = hub_rows[2 * i + 1] >> 20;
        const uint32_t rs = row[u], re = row[u + 1];
        const uint32_t e0 = rs + c * AGG_HUB, e1 = min(re, e0 + AGG_HUB);
        float s0 = 0.f, s1 = 0.f;
        for (uint32_t e = e0 + slot; e < e1; e += SLOTS) {
            const uint32_t nb = (uint32_t)col[e];
            if (BWD) {
                const float* in = (const float*)in_;
                const float sc = 1.f / (float)(row[nb + 1] - row[nb]);
                s0 += in[(uint64_t)nb * D + 2 * fp] * sc;
                s1 += in[(uint64_t)nb * D + 2 * fp + 1] * sc;
            } else {
                const bf16* in = (const bf16*)in_;
                s0 += bf2f(in[(uint64_t)nb * D + 2 * fp]);
                s1 += bf2f(in[(uint64_t)nb * D + 2 * fp + 1]);
            }
        }
#pragma unroll
        for (int o = FP; o < 64; o <<= 1) {
            s0 += __shfl_xor(s0, o, 64);
            s1 += __shfl_xor(s1, o, 64);
        }
        if (slot == 0) {
            atomicAdd(&hub_acc[(uint64_t)h * D + 2 * fp], s0);
            atomicAdd(&hub_acc[(uint64_t)h * D + 2 * fp + 1], s1);
        }
    }
}

// finish hub rows: forward writes mean (bf16), backward adds the sums
template <int D, bool BWD>
__global__ void k_aggregate_hub_finish(const uint32_t* row, const uint32_t* hub_list, uint32_t n_hub, const float* hub_acc,
                                       void* out_) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_hub * D; i += gridDim.x * blockDim.x) {
        const uint32_t h = i / D, f = i % D;
        const uint32_t u = hub_list[h];
        if (BWD) {
            ((float*)out_)[(uint64_t)u * D + f] += hub_acc[i];
        } else {
            const float deg = (float)(row[u + 1] - row[u]);
            ((bf16*)out_)[(uint64_t)u * D + f] = f2bf(hub_acc[i] / deg);
        }
    }
}

// ---------------------------------------------------------------------------------------
// dense layers on MFMA (v_mfma_f32_16x16x32_bf16).  Lane l of a 16x16x32 step holds
//   A[l%16][8(l/16)+j], B[8(l/16)+j][l%16] (j<8) and D[4(l/16)+i][l%16] (i<4).
// ---------------------------------------------------------------------------------------

// out[v][n] = relu(sum_k X[v][k] Wt[n][k] + b[n]), X = [self | nbr] (D each), K = 2D, N = 64.
// One wave per 16-row tile.  HEAD: also the score s[v] = b3 + sum_n w3[n] out[v][n] (the bf16
// activations, summed over each row's 16 lanes by shuffles), and out may be nullptr (inference).
template <int D, bool HEAD>
__global__ void __launch_bounds__(256) k_dense_fwd(const bf16* self, const bf16* nbr, const bf16* Wt, const float* b,
                                                   uint32_t n, bf16* out, const float* w3, const float* b3, float* s) {
    constexpr int K = 2 * D;
    const uint32_t lane = __lane_id();
    const uint32_t r = lane % 16, q = lane / 16;
    const uint32_t gwave = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const uint32_t nwaves = gridDim.x * blockDim.x / 64;
    const uint32_t ntiles = (n + 15) / 16;
    // the wave's B fragments (all of W, K/32 x 4 steps of 8 bf16 per lane) and biases stay in
    // registers across its row tiles
    constexpr int KS = K / 32;
    bf16x8 bw[KS][4];
    float bias[4];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) bw[ks][nb] = *(const bf16x8*)(Wt + (uint64_t)(nb * 16 + r) * K + ks * 32 + 8 * q);
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) bias[nb] = b[nb * 16 + r];
    float w3v[4] = {0.f, 0.f, 0.f, 0.f};
    float b3v = 0.f;
    if (HEAD) {
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) w3v[nb] = w3[nb * 16 + r];
        b3v = b3[0];
    }
    for (uint32_t t = gwave; t < ntiles; t += nwaves) {
        const uint32_t v0 = t * 16;
        floatx4 acc[4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
        const uint32_t va = v0 + r;
        bf16x8 a[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int kk = ks * 32 + 8 * (int)q;   // this lane's 8 consecutive k
            if (va < n) {
                const bf16* src = kk < D ? self + (uint64_t)va * D + kk : nbr + (uint64_t)va * D + (kk - D);
                a[ks] = *(const bf16x8*)src;
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) a[ks][j] = f2bf(0.f);
            }
        }
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ks], bw[ks][nb], acc[nb], 0, 0, 0);
        if (HEAD) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t v = v0 + 4 * q + i;
                float part = 0.f;
#pragma unroll
                for (int nb = 0; nb < 4; ++nb) {
                    const bf16 h = f2bf(fmaxf(acc[nb][i] + bias[nb], 0.f));
                    if (out && v < n) out[(uint64_t)v * HID + nb * 16 + r] = h;
                    part += w3v[nb] * bf2f(h);
                }
#pragma unroll
                for (int o = 1; o < 16; o <<= 1) part += __shfl_xor(part, o, 64);   // the row's 16 lanes
                if (r == 0 && v < n) s[v] = b3v + part;
            }
        } else {
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                const uint32_t col_n = nb * 16 + r;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint32_t v = v0 + 4 * q + i;
                    if (v < n) out[(uint64_t)v * HID + col_n] = f2bf(fmaxf(acc[nb][i] + bias[nb], 0.f));
                }
            }
        }
    }
}

// dX[v][k] = sum_n dZ[v][n] W[n][k]  (dZ bf16 [n x 64], Wkn bf16 [K][64] = W transposed, dX fp32)
template <int K>
__global__ void __launch_bounds__(256) k_dense_bwd_x(const bf16* dZ, const bf16* Wkn, uint32_t n, float* dX) {
    const uint32_t lane = __lane_id();
    const uint32_t r = lane % 16, q = lane / 16;
    const uint32_t gwave = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const uint32_t nwaves = gridDim.x * blockDim.x / 64;
    const uint32_t ntiles = (n + 15) / 16;
    for (uint32_t t = gwave; t < ntiles; t += nwaves) {
        const uint32_t v0 = t * 16;
        const uint32_t va = v0 + r;
        bf16x8 a0, a1;   // dZ row chunk: n in [8q, 8q+8) and [32+8q, 32+8q+8)
        if (va < n) {
            a0 = *(const bf16x8*)(dZ + (uint64_t)va * HID + 8 * q);
            a1 = *(const bf16x8*)(dZ + (uint64_t)va * HID + 32 + 8 * q);
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) { a0[j] = f2bf(0.f); a1[j] = f2bf(0.f); }
        }
#pragma unroll
        for (int kb = 0; kb < K / 16; ++kb) {
            floatx4 acc = {0, 0, 0, 0};
            // B[n][k] = W[n][k] = Wkn[k][n]: lane needs n = 8q..8q+7 at k = kb*16 + r
            const bf16x8 b0 = *(const bf16x8*)(Wkn + (uint64_t)(kb * 16 + r) * HID + 8 * q);
            const bf16x8 b1 = *(const bf16x8*)(Wkn + (uint64_t)(kb * 16 + r) * HID + 32 + 8 * q);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t v = v0 + 4 * q + i;
                if (v < n) dX[(uint64_t)v * K + kb * 16 + r] = acc[i];
            }
        }
    }
}

// dW[n][k] += sum_v dZ[v][n] X[v][k], X = [self | nbr]; reduction over v in 32-row chunks staged
// in LDS (transposed operands), partial sums of a workgroup added atomically.
template <int D>
__global__ void __launch_bounds__(256) k_dense_bwd_w(const bf16* dZ, const bf16* self, const bf16* nbr, uint32_t n,
                                                     float* dW) {
    constexpr int K = 2 * D;
    constexpr int NT = 4 * (K / 16);           // output tiles (n-block, k-block)
    constexpr int TPW = (NT + 3) / 4;          // tiles per wave
    __shared__ bf16 sZ[32][HID + 8];
    __shared__ bf16 sX[32][K + 8];
    const uint32_t lane = __lane_id(), wid = threadIdx.x / 64;
    const uint32_t r = lane % 16, q = lane / 16;
    floatx4 acc[TPW];
#pragma unroll
    for (int t = 0; t < TPW; ++t) acc[t] = {0, 0, 0, 0};
    const uint32_t nchunks = (n + 31) / 32;
    for (uint32_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const uint32_t v0 = ch * 32;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < 32 * HID; i += blockDim.x) {
            const uint32_t vv = i / HID, f = i % HID;
            sZ[vv][f] = (v0 + vv < n) ? dZ[(uint64_t)(v0 + vv) * HID + f] : f2bf(0.f);
        }
        for (uint32_t i = threadIdx.x; i < 32 * K; i += blockDim.x) {
            const uint32_t vv = i / K, k = i % K;
            bf16 x = f2bf(0.f);
            if (v0 + vv < n) x = k < (uint32_t)D ? self[(uint64_t)(v0 + vv) * D + k] : nbr[(uint64_t)(v0 + vv) * D + (k - D)];
            sX[vv][k] = x;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < TPW; ++t) {
            const int tile = (int)wid * TPW + t;
            if (tile >= NT) break;
            const int nb = tile % 4, kb = tile / 4;
            bf16x8 a, bv;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                a[j] = sZ[8 * q + j][nb * 16 + r];     // A[n][vv] = dZ[vv][n]
                bv[j] = sX[8 * q + j][kb * 16 + r];    // B[vv][k] = X[vv][k]
            }
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        const int tile = (int)wid * TPW + t;
        if (tile >= NT) break;
        const int nb = tile % 4, kb = tile / 4;
#pragma unroll
        for (int i = 0; i < 4; ++i) atomicAdd(&dW[(uint64_t)(nb * 16 + 4 * q + i) * K + kb * 16 + r], acc[t][i]);
    }
}

// s_v = w3 . h2_v + b3
// weighted logistic loss: L = sum_v wt_v * (softplus(s) - y s); dS = wt_v * (sigmoid(s) - y)
// wt_v = pos_weight for y=1, 1 otherwise; everything divided by n.
__global__ void k_loss(const float* s, const uint8_t* y, uint32_t n, float pos_weight, float* dS, float* loss) {
    float l = 0.f;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const float sv = s[v];
        const float yv = y[v] ? 1.f : 0.f;
        const float wt = (y[v] ? pos_weight : 1.f) / (float)n;
        const float sp = sv > 0 ? sv + log1pf(expf(-sv)) : log1pf(expf(sv));
        l += wt * (sp - yv * sv);
        const float sig = 1.f / (1.f + expf(-sv));
        dS[v] = wt * (sig - yv);
    }
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
    if (__lane_id() == 0) atomicAdd(loss, l);
}

// head backward: grad w3 += sum dS h2, grad b3 += sum dS, dZ2 = dS w3 (relu mask), as bf16
__global__ void k_head_bwd(const float* dS, const bf16* h2, const float* w3, uint32_t n, bf16* dZ2, float* gw3, float* gb3) {
    __shared__ float sw[HID + 1];
    if (threadIdx.x <= HID) sw[threadIdx.x] = 0.f;
    __syncthreads();
    const uint32_t f = threadIdx.x % HID;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (uint64_t)n * HID;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t v = (uint32_t)(i / HID);
        const float h = bf2f(h2[i]);
        const float d = dS[v];
        atomicAdd(&sw[f], d * h);
        if (f == 0) atomicAdd(&sw[HID], d);
        dZ2[i] = f2bf(h > 0.f ? d * w3[f] : 0.f);
    }
    __syncthreads();
    if (threadIdx.x < HID) atomicAdd(&gw3[threadIdx.x], sw[threadIdx.x]);
    if (threadIdx.x == HID) atomicAdd(gb3, sw[HID]);
}

// layer-2 input gradient split: dH1 = dX2[:, :64] (self part); dA1 = dX2[:, 64:] goes through the
// transposed aggregation; then dZ1 = dH1 (relu mask of h1), bias grads
__global__ void k_split_self(const float* dX2, uint32_t n, float* dH1, float* dA1) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (uint64_t)n * HID;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t v = i / HID, f = i % HID;
        dH1[i] = dX2[v * K2 + f];
        dA1[i] = dX2[v * K2 + HID + f];
    }
}

__global__ void k_relu_mask_bias(const float* dH, const bf16* h, uint32_t n, bf16* dZ, float* gb) {
    __shared__ float sb[HID];
    if (threadIdx.x < HID) sb[threadIdx.x] = 0.f;
    __syncthreads();
    const uint32_t f = threadIdx.x % HID;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (uint64_t)n * HID;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const float d = bf2f(h[i]) > 0.f ? dH[i] : 0.f;
        const bf16 db = f2bf(d);
        dZ[i] = db;
        atomicAdd(&sb[f], bf2f(db));
    }
    __syncthreads();
    if (threadIdx.x < HID) atomicAdd(&gb[threadIdx.x], sb[threadIdx.x]);
}

// layer-2 bias gradient from dZ2 (bf16)
__global__ void k_bias_grad(const bf16* dZ, uint32_t n, float* gb) {
    __shared__ float sb[HID];
    if (threadIdx.x < HID) sb[threadIdx.x] = 0.f;
    __syncthreads();
    const uint32_t f = threadIdx.x % HID;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (uint64_t)n * HID;
         i += (uint64_t)gridDim.x * blockDim.x)
        atomicAdd(&sb[f], bf2f(dZ[i]));
    __syncthreads();
    if (threadIdx.x < HID) atomicAdd(&gb[threadIdx.x], sb[threadIdx.x]);
}

// Adam on fp32 master parameters; refresh the bf16 copies (Wt row-major and transposed)
__global__ void k_adam(float* p, float* g, float* m, float* v, int np, float lr, float b1, float b2, float eps,
                       float bc1, float bc2, float inv_count, bf16* W1t, bf16* W2t, bf16* W1kn, bf16* W2kn) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x) {
        const float gi = g[i] * inv_count;
        m[i] = b1 * m[i] + (1.f - b1) * gi;
        v[i] = b2 * v[i] + (1.f - b2) * gi * gi;
        const float mh = m[i] / bc1, vh = v[i] / bc2;
        p[i] -= lr * mh / (sqrtf(vh) + eps);
        g[i] = 0.f;
    }
}

__global__ void k_refresh_bf16(const float* p, bf16* W1t, bf16* W2t, bf16* W1kn, bf16* W2kn) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < HID * K2; i += gridDim.x * blockDim.x) {
        if (i < HID * K1) {
            const int nn = i / K1, k = i % K1;
            W1t[i] = f2bf(p[OFF_W1 + i]);
            W1kn[k * HID + nn] = f2bf(p[OFF_W1 + i]);
        }
        const int nn = i / K2, k = i % K2;
        W2t[i] = f2bf(p[OFF_W2 + i]);
        W2kn[k * HID + nn] = f2bf(p[OFF_W2 + i]);
    }
}

// ---------------------------------------------------------------------------------------
// top-k selection and induced subgraph
// ---------------------------------------------------------------------------------------
// order-preserving key of a float score (larger score -> larger key)
__device__ __forceinline__ uint32_t score_key(float s) {
    const uint32_t b = __float_as_uint(s);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// Top-k threshold by a three-pass radix select (11, 11 and 10 bits, high to low) over
// order-preserving keys, with no host round trip.  sel_state: [0] keys still wanted (k at the
// start, the tie budget at the end), [1] the digits of the threshold key found so far.
__device__ __forceinline__ void radix_digit(int pass, uint32_t& lo, uint32_t& nb) {
    lo = pass == 0 ? 21u : (pass == 1 ? 10u : 0u);
    nb = pass == 2 ? 10u : 11u;
}

// histogram of this pass's digit over the keys whose higher bits equal the digits found so far,
// privatised in LDS per workgroup (only non-empty bins reach the global histogram)
__global__ void __launch_bounds__(256) k_key_hist(const float* s, uint32_t n, int pass, const uint32_t* sel_state,
                                                  uint32_t* hist) {
    __shared__ uint32_t lh[2048];
    uint32_t lo, nb;
    radix_digit(pass, lo, nb);
    const uint32_t hi = lo + nb;
    const uint32_t prefix = sel_state[1];
    const uint32_t nbins = 1u << nb;
    for (uint32_t i = threadIdx.x; i < nbins; i += blockDim.x) lh[i] = 0;
    __syncthreads();
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const uint32_t k = score_key(s[v]);
        if (hi >= 32u || (k >> hi) == prefix) atomicAdd(&lh[(k >> lo) & (nbins - 1u)], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nbins; i += blockDim.x)
        if (lh[i]) atomicAdd(&hist[i], lh[i]);
}

// the digit whose bin, counted from the top, first reaches the wanted number (1 <= wanted <= keys
// left): thread t sums nbins/256 bins from the top, a prefix over the 256 partial sums finds the
// thread whose range holds the crossing, and that thread walks its bins and appends the digit.
// Each thread then clears its own bins, so the next pass (or selection) starts from a zero histogram.
// The first pass starts the state itself: k_target keys wanted, no digits yet.
__global__ void __launch_bounds__(256) k_pick_bin(uint32_t* hist, uint32_t* sel_state, int pass, uint32_t k_target) {
    __shared__ uint32_t part[256];
    uint32_t lo, nb;
    radix_digit(pass, lo, nb);
    const uint32_t nbins = 1u << nb, per = nbins / 256u;
    const uint32_t t = threadIdx.x;
    const uint32_t need = pass == 0 ? k_target : sel_state[0];
    const uint32_t prefix = pass == 0 ? 0u : sel_state[1];   // read before the barrier below
    const uint32_t top = nbins - 1u - per * t;
    uint32_t sum = 0;
    for (uint32_t j = 0; j < per; ++j) sum += hist[top - j];
    part[t] = sum;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t j = 0; j < t; ++j) before += part[j];
    if (before < need && before + sum >= need) {
        uint32_t acc = before;
        for (uint32_t j = 0; j < per; ++j) {
            const uint32_t h = hist[top - j];
            if (acc + h >= need) {
                sel_state[1] = (prefix << nb) | (top - j);
                sel_state[0] = need - acc;
                break;
            }
            acc += h;
        }
    }
    for (uint32_t j = 0; j < per; ++j) hist[top - j] = 0u;
}

// per-block exclusive scan of a u32 array (block = 1024 elements), block sums out
__global__ void __launch_bounds__(256) k_scan_blocks(const uint32_t* in, uint32_t n, uint32_t* out, uint32_t* bsum) {
    __shared__ uint32_t sw[4];
    const uint32_t base = blockIdx.x * 1024;
    const uint32_t lane = __lane_id(), wid = threadIdx.x / 64;
    uint32_t vals[4], tot = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t i = base + threadIdx.x * 4 + j;
        vals[j] = i < n ? in[i] : 0u;
        tot += vals[j];
    }
    uint32_t incl = tot;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(incl, d, 64);
        if (lane >= (uint32_t)d) incl += y;
    }
    if (lane == 63) sw[wid] = incl;
    __syncthreads();
    uint32_t off = 0;
    for (uint32_t w2 = 0; w2 < wid; ++w2) off += sw[w2];
    uint32_t run = off + incl - tot;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t i = base + threadIdx.x * 4 + j;
        if (i < n) out[i] = run;
        run += vals[j];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = sw[0] + sw[1] + sw[2] + sw[3];
}

// the selection and the first pass of its scan in one: v is selected when key > threshold or key ==
// threshold within the tie budget, plus every group member; sel[v] out, then as k_scan_blocks
// (exclusive per-block offsets of the selected vertices, block sums out)
__global__ void __launch_bounds__(256) k_select_scan(const float* s, const uint32_t* gm, uint32_t n,
                                                     uint32_t* sel_state, uint32_t* sel, uint32_t* out,
                                                     uint32_t* bsum) {
    __shared__ uint32_t sw[4];
    const uint32_t thr = sel_state[1];
    int* tie_budget = (int*)&sel_state[0];
    const uint32_t base = blockIdx.x * 1024;
    const uint32_t lane = __lane_id(), wid = threadIdx.x / 64;
    uint32_t vals[4], tot = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t v = base + threadIdx.x * 4 + j;
        uint32_t x = 0;
        if (v < n) {
            const uint32_t k = score_key(s[v]);
            x = (k > thr) ? 1u : 0u;
            if (k == thr && atomicSub(tie_budget, 1) > 0) x = 1u;   // signed: stays <= 0 once spent
            if (gm[v]) x = 1u;
            sel[v] = x;
        }
        vals[j] = x;
        tot += x;
    }
    uint32_t incl = tot;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(incl, d, 64);
        if (lane >= (uint32_t)d) incl += y;
    }
    if (lane == 63) sw[wid] = incl;
    __syncthreads();
    uint32_t off = 0;
    for (uint32_t w2 = 0; w2 < wid; ++w2) off += sw[w2];
    uint32_t run = off + incl - tot;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t i = base + threadIdx.x * 4 + j;
        if (i < n) out[i] = run;
        run += vals[j];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = sw[0] + sw[1] + sw[2] + sw[3];
}

// exclusive scan of the block sums by one wave (64 at a time with a running offset), total in
// bsum[nb] and, when given, in *total_out
__global__ void __launch_bounds__(64) k_scan_sums(uint32_t* bsum, uint32_t nb, uint32_t* total_out) {
    if (blockIdx.x != 0) return;
    const uint32_t lane = threadIdx.x;
    uint32_t run = 0;
    for (uint32_t base = 0; base < nb; base += 64) {
        const uint32_t i = base + lane;
        const uint32_t x = i < nb ? bsum[i] : 0u;
        uint32_t incl = x;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(incl, d, 64);
            if (lane >= (uint32_t)d) incl += y;
        }
        if (i < nb) bsum[i] = run + incl - x;
        run += __shfl(incl, 63, 64);
    }
    if (lane == 0) {
        bsum[nb] = run;
        if (total_out) *total_out = run;
    }
}


// the last pass of the new-id scan (block offsets added), the list of the selected vertices by new
// id (old_of[newid[v]] = v), and every degree cleared
__global__ void k_scan_add_list(uint32_t* newid, uint32_t n, const uint32_t* bsum, const uint32_t* sel,
                                uint32_t* old_of, uint32_t* deg) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const uint32_t id = newid[v] + bsum[v / 1024];
        newid[v] = id;
        deg[v] = 0u;
        if (sel[v]) old_of[id] = v;
    }
}

// degree of each selected vertex inside the selection (one wave per selected vertex; *ns of them)
// The group members' new ids (gv: cnt member entries) are written first, one thread per entry.
__global__ void k_sub_deg(const uint32_t* row, const int32_t* col, const uint32_t* sel, const uint32_t* old_of,
                          const uint32_t* ns, uint32_t* deg, const int32_t* gv, uint32_t gcnt, const uint32_t* newid,
                          int32_t* sub_gv) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < gcnt; i += gridDim.x * blockDim.x)
        sub_gv[i] = (int32_t)newid[gv[i]];
    const uint32_t lane = __lane_id();
    const uint32_t gwave = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const uint32_t nwaves = gridDim.x * blockDim.x / 64;
    const uint32_t cnt = *ns;
    for (uint32_t i = gwave; i < cnt; i += nwaves) {
        const uint32_t v = old_of[i];
        const uint32_t re = row[v + 1];
        uint32_t c = 0;
        // 4 x 64 arcs per step: their column loads, then their selection gathers, all independent
        // (a hub row of thousands of arcs is one wave's work)
        for (uint32_t e0 = row[v]; e0 < re; e0 += 256) {
            uint32_t cc[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t e = e0 + j * 64 + lane;
                cc[j] = e < re ? (uint32_t)col[e] : 0xFFFFFFFFu;
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) c += cc[j] != 0xFFFFFFFFu ? sel[cc[j]] : 0u;
        }
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
        if (lane == 0) deg[v] = c;
    }
}

// fill the subgraph CSR: new row i holds old vertex old_of[i]; arcs keep their order (new ids are
// monotone in the old ones, so rows stay in (weight, neighbour) order); each arc is one record
// (weight bits << 32 | new neighbour id), as the solver reads them
// arc_off: per-block exclusive offsets of the arc scan, arc_bsum: its scanned block sums (added here,
// as the scan's last pass), total: the arc count (the subgraph's closing row offset)
__global__ void k_sub_fill(const uint32_t* row, const uint64_t* arc, const uint32_t* sel, const uint32_t* newid,
                           const uint32_t* arc_off, const uint32_t* arc_bsum, const uint32_t* total,
                           const uint32_t* old_of, uint32_t ns, uint32_t* srow, uint64_t* sarc) {
    const uint32_t lane = __lane_id();
    const uint32_t gwave = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const uint32_t nwaves = gridDim.x * blockDim.x / 64;
    if (blockIdx.x == 0 && threadIdx.x == 0) srow[ns] = *total;
    for (uint32_t nv = gwave; nv < ns; nv += nwaves) {
        const uint32_t v = old_of[nv];
        uint32_t pos = arc_off[v] + arc_bsum[v / 1024];
        if (lane == 0) srow[nv] = pos;
        const uint32_t re = row[v + 1];
        // 4 x 64 arcs per step: their records, then their selection flags and new ids, all
        // independent loads; the positions follow from the ballots in arc order
        for (uint32_t e0 = row[v]; e0 < re; e0 += 256) {
            uint64_t r[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t e = e0 + j * 64 + lane;
                r[j] = e < re ? arc[e] : ~0ull;
            }
            uint32_t sl[4], nid[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t x = (uint32_t)r[j];
                const bool in = r[j] != ~0ull;
                sl[j] = in ? sel[x] : 0u;
                nid[j] = in ? newid[x] : 0u;
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool ok = sl[j] != 0u;
                const uint64_t bal = __ballot(ok);
                if (ok) {
                    const uint32_t p = pos + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
                    sarc[p] = (r[j] & 0xFFFFFFFF00000000ull) | nid[j];
                }
                pos += (uint32_t)__popcll(bal);
            }
        }
    }
}


// Layer-1 aggregation of a query splits into a graph-only part and a group part: a vertex's input
// features are two graph-only values (degree, lightest arc), the membership bits of groups 0..7 and
// zeros, so the neighbour mean of the graph-only values is computed once per graph (ax_static,
// with the gather kernel on features whose bits are all clear) and the mean of bit i is the count
// of the row's neighbours in group i over the degree.  The counts come from a scatter along the
// members' rows: one wave per member entry sets the member's group bit (for every group, as the
// solver's masks need) and, the first time a (member, group < 8) pair is seen, adds 1 to the
// fp32 count of each neighbour (integers, so exact in any order).
__global__ void __launch_bounds__(256) k_set_bits_scatter(uint32_t* gm, const int32_t* ptr, const int32_t* vtx, int g,
                                                          const uint32_t* row, const int32_t* col, float* gcnt) {
    const uint32_t lane = __lane_id();
    const uint32_t gwave = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const uint32_t nwaves = gridDim.x * blockDim.x / 64;
    const uint32_t tot = (uint32_t)ptr[g];
    for (uint32_t j = gwave; j < tot; j += nwaves) {
        int i = 0;
        while (i + 1 < g && (uint32_t)ptr[i + 1] <= j) ++i;
        const uint32_t u = (uint32_t)vtx[j];
        uint32_t old = 0;
        if (lane == 0) old = atomicOr(&gm[u], 1u << i);
        old = __shfl(old, 0, 64);
        if (i >= 8 || ((old >> i) & 1u)) continue;
        for (uint32_t e = row[u] + lane; e < row[u + 1]; e += 64) atomicAdd(&gcnt[(uint64_t)col[e] * 8 + i], 1.0f);
    }
}

// ax[v] = the graph-only means of ax_static with the group means count / degree, rounded as the
// gather rounds them (sum times 1/degree on light rows, sum / degree on hub rows); the counts are
// cleared behind the read, ready for the next query's scatter
__global__ void k_ax_finish(const uint32_t* row, const bf16* ax_static, float* gcnt, uint32_t n, bf16* ax) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const uint32_t deg = row[v + 1] - row[v];
        const float inv = deg ? 1.f / (float)deg : 0.f;
        bf16x8 a = *(const bf16x8*)(ax_static + (uint64_t)v * FIN);
        bf16x8 b;
#pragma unroll
        for (int j = 0; j < 8; ++j) b[j] = f2bf(0.f);
        const float4 c0 = *(const float4*)(gcnt + (uint64_t)v * 8);
        const float4 c1 = *(const float4*)(gcnt + (uint64_t)v * 8 + 4);
        const float c[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
        *(float4*)(gcnt + (uint64_t)v * 8) = make_float4(0.f, 0.f, 0.f, 0.f);
        *(float4*)(gcnt + (uint64_t)v * 8 + 4) = make_float4(0.f, 0.f, 0.f, 0.f);
        float m[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) m[i] = deg > (uint32_t)AGG_HUB ? c[i] / (float)deg : c[i] * inv;
#pragma unroll
        for (int i = 0; i < 6; ++i) a[2 + i] = f2bf(m[i]);
        b[0] = f2bf(m[6]);
        b[1] = f2bf(m[7]);
        bf16x8* o = (bf16x8*)(ax + (uint64_t)v * FIN);
        o[0] = a;
        o[1] = b;
    }
}

__global__ void k_fill32(uint32_t* p, uint32_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

int gridn(uint64_t work, int block, int cap) {
    uint64_t b = (work + block - 1) / block;
    if (b < 1) b = 1;
    if (b > (uint64_t)cap) b = cap;
    return (int)b;
}

template <class T>
bool dalloc(T*& p, size_t elems) {
    hipFree(p);
    p = nullptr;
    return hipMalloc((void**)&p, std::max<size_t>(elems, 1) * sizeof(T)) == hipSuccess;
}

}  // namespace

// =======================================================================================
// One selection in flight (scores, top-k, induced subgraph of one query): its own scratch, pinned
// staging and events, so that several selections can be queued on one stream without the host
// waiting between them (they share the model's activations in stream order).
struct SelSlot {
    uint32_t* scr = nullptr;     // device: hist 2048 | state 8 | sel n | newid n | deg n | arcoff n | old n |
    uint64_t cap = 0;            //         block sums 2 (nbk + 2) | ns, ms, sub group ids gcnt
    uint32_t* h_rb = nullptr;    // pinned readback: ns, ms, sub group ids
    uint32_t h_cap = 0;
    int32_t* h_grp = nullptr;    // pinned group lists of the query (upload source)
    uint32_t hg_cap = 0;
    hipEvent_t e_sizes = nullptr, e_built = nullptr;
    uint32_t n = 0, nbk = 0, gcnt = 0;
    hipStream_t st = nullptr;    // stream of the selection in flight (its activation set's)
};

void sel_free(SelSlot& sl) {
    hipFree(sl.scr);
    if (sl.h_rb) hipHostFree(sl.h_rb);
    if (sl.h_grp) hipHostFree(sl.h_grp);
    if (sl.e_sizes) hipEventDestroy(sl.e_sizes);
    if (sl.e_built) hipEventDestroy(sl.e_built);
    sl = SelSlot();
}

// device CSR of one induced subgraph (reused across queries and calls while large enough)
struct SubBuf {
    uint32_t* row = nullptr;
    uint64_t* arc = nullptr;   // arc records (weight bits << 32 | neighbour)
    uint32_t cap_n = 0;
    uint64_t cap_m = 0;
};

static void sub_free(SubBuf& b) {
    hipFree(b.row); hipFree(b.arc);
    b = SubBuf();
}

// room for ns vertices and ms arcs; a new allocation takes a quarter more, and at least
// floor_n / floor_m (a pool passes the largest sizes it has seen, so that its buffers rarely
// grow again: a free waits for the whole device)
static bool sub_reserve(SubBuf& b, uint32_t ns, uint32_t ms, uint32_t floor_n = 0, uint64_t floor_m = 0) {
    if (b.cap_n >= ns + 1 && b.cap_m >= std::max(ms, 1u)) return true;
    sub_free(b);
    const uint32_t cn = std::max<uint32_t>(ns + 1 + ns / 4 + 64, floor_n);
    const uint64_t cm = std::max<uint64_t>((uint64_t)std::max(ms, 1u) + ms / 4 + 256, floor_m);
    if (hipMalloc(&b.row, cn * 4ull) != hipSuccess || hipMalloc(&b.arc, cm * 8) != hipSuccess) {
        sub_free(b);
        return false;
    }
    b.cap_n = cn;
    b.cap_m = cm;
    return true;
}

// Activations of one forward pass.  The batched learned path runs two sets, each on a stream of its
// own, so that the selections of two queries overlap; training and the one-query path use set 0.
struct FwdSet {
    bf16 *x = nullptr, *ax = nullptr, *h1 = nullptr, *ah1 = nullptr;
    float* s = nullptr;          // scores
    uint32_t* gm = nullptr;      // group bits per vertex
    float* gcnt = nullptr;       // per vertex, neighbours in each of groups 0..7 (cleared after reading)
    float* hub_acc = nullptr;    // hub-row partial sums of an aggregation
    int32_t* grp = nullptr;      // the query's group lists
    uint32_t cap_n = 0, cap_grp = 0, cap_hub = 0;
    hipStream_t st = nullptr;    // nullptr: the model's stream
};
constexpr int FWD_SETS = 2;

struct gst_model {
    gst_ctx* ctx = nullptr;
    float* p = nullptr;   // NPARAM fp32 master
    float* g = nullptr;   // gradient accumulator
    float* m = nullptr;
    float* v = nullptr;
    bf16* W1t = nullptr;  // [HID][K1]
    bf16* W2t = nullptr;  // [HID][K2]
    bf16* W1kn = nullptr; // [K1][HID]
    bf16* W2kn = nullptr; // [K2][HID]
    int steps = 0;
    int grad_count = 0;
    // forward activations (set 0: training and the one-query path; both: batched selections)
    FwdSet f[FWD_SETS];
    // training buffers and the per-graph part of layer 1 (grown on demand)
    uint32_t cap_n = 0;
    bf16 *h2 = nullptr, *dZ2 = nullptr, *dZ1 = nullptr;
    float *dS = nullptr, *dX2 = nullptr, *dH1 = nullptr, *dA1 = nullptr, *loss = nullptr;
    bf16* ax_static = nullptr;   // graph-only part of the layer-1 aggregation (per graph)
    uint64_t static_uid = 0;     // graph ax_static belongs to
    hipEvent_t e_static = nullptr;   // ax_static written (other streams wait for it)
    uint8_t* y = nullptr;
    // hub rows of the last graph
    uint64_t hub_uid = 0;
    SelSlot sync_sel;                  // the selection of the one-query path
    std::vector<SelSlot> batch_sel;    // selections in flight of the batched path
    SubBuf sync_buf;                   // subgraph CSR of the one-query path, kept across calls
    std::vector<SubBuf> sub_pool;      // subgraph CSRs of the batched path, kept across calls
    uint32_t sub_max_n = 0;            // largest subgraph rows / arcs a pool buffer was sized for
    uint64_t sub_max_m = 0;
    void* sel_stream = nullptr;        // stream of the forward pass and selection (nullptr: the context's)
    uint32_t* hub_list = nullptr;   // hub row ids
    uint32_t* hub_items = nullptr;  // pairs (row, h | chunk << 20)
    uint32_t n_hub = 0, n_hub_items = 0;
    // graphs whose selections gave oversized subgraphs: (graph uid, how many); set and read by one
    // host thread at a time (the caller's, or a batched call's selection thread)
    std::vector<std::pair<uint64_t, int>> oversized;
};

namespace {

hipStream_t mstream(const gst_model* M) {
    return (hipStream_t)(M->sel_stream ? M->sel_stream : M->ctx->stream);
}

// a selection of ms arcs on G: is the subgraph too large to solve?  (counted per graph)
bool sub_oversized(gst_model* M, const gst_graph* G, uint64_t ms) {
    if (ms * 1000ull <= (uint64_t)G->m * GST_SUB_ARC_PERMILLE) return false;
    for (auto& p : M->oversized)
        if (p.first == G->uid) { p.second++; return true; }
    if (M->oversized.size() >= 64) M->oversized.erase(M->oversized.begin());
    M->oversized.push_back({G->uid, 1});
    return true;
}

// has the model stopped selecting on G (two oversized subgraphs)?
bool sub_declined(const gst_model* M, const gst_graph* G) {
    for (auto& p : M->oversized)
        if (p.first == G->uid) return p.second >= 2;
    return false;
}

template <int D, bool BWD>
void aggregate(gst_model* M, const gst_graph* G, float* hub_acc, const void* in, void* out, hipStream_t st);

hipStream_t fstream(const gst_model* M, const FwdSet& F) { return F.st ? F.st : mstream(M); }

// one activation set's buffers for an n-vertex graph with n_hub hub rows (its group counts start clear;
// every forward pass clears them again after reading)
gst_status fwd_bufs(gst_model* M, FwdSet& F, uint32_t n) {
    gst_ctx* ctx = M->ctx;
    if (F.cap_n < n) {
        const bool ok = dalloc(F.x, (size_t)n * FIN) && dalloc(F.ax, (size_t)n * FIN) && dalloc(F.h1, (size_t)n * HID) &&
                        dalloc(F.ah1, (size_t)n * HID) && dalloc(F.s, n) && dalloc(F.gm, n) && dalloc(F.gcnt, (size_t)n * 8);
        if (!ok) { F.cap_n = 0; return gst_set_error(ctx, GST_OOM, "model: activation allocation failed (n=%u)", n); }
        F.cap_n = n;
        const int fgz = std::max(1, ctx->cu_count * 8);
        k_fill32<<<gridn((uint64_t)n * 8, 256, fgz), 256, 0, fstream(M, F)>>>((uint32_t*)F.gcnt, 0u, (uint64_t)n * 8);
    }
    if (F.cap_hub < std::max(M->n_hub, 1u)) {
        if (!dalloc(F.hub_acc, (size_t)std::max(M->n_hub, 1u) * HID)) {
            F.cap_hub = 0;
            return gst_set_error(ctx, GST_OOM, "model: hub allocation failed");
        }
        F.cap_hub = std::max(M->n_hub, 1u);
    }
    return GST_OK;
}

gst_status model_bufs(gst_model* M, const gst_graph* G) {
    gst_ctx* ctx = M->ctx;
    const uint32_t n = (uint32_t)G->n;
    if (M->cap_n < n) {
        bool ok = dalloc(M->h2, (size_t)n * HID) && dalloc(M->dZ2, (size_t)n * HID) && dalloc(M->dZ1, (size_t)n * HID) &&
                  dalloc(M->dS, n) && dalloc(M->dX2, (size_t)n * K2) && dalloc(M->dH1, (size_t)n * HID) &&
                  dalloc(M->dA1, (size_t)n * HID) && dalloc(M->y, n) && dalloc(M->ax_static, (size_t)n * FIN);
        M->static_uid = 0;
        if (!ok) { M->cap_n = 0; return gst_set_error(ctx, GST_OOM, "model: activation allocation failed (n=%u)", n); }
        M->cap_n = n;
    }
    if (M->hub_uid != G->uid) {
        // hub rows (host side, from the graph's host CSR)
        std::vector<uint32_t> hl, hi;
        for (uint32_t v = 0; v < n; ++v) {
            const int64_t d = G->h_row[v + 1] - G->h_row[v];
            if (d > AGG_HUB) {
                const uint32_t h = (uint32_t)hl.size();
                hl.push_back(v);
                for (int64_t c = 0; c * AGG_HUB < d; ++c) { hi.push_back(v); hi.push_back(h | ((uint32_t)c << 20)); }
            }
        }
        if (hl.size() >= (1u << 20)) return gst_set_error(ctx, GST_INVALID_ARG, "model: too many hub rows");
        if (!dalloc(M->hub_list, hl.size()) || !dalloc(M->hub_items, hi.size()))
            return gst_set_error(ctx, GST_OOM, "model: hub allocation failed");
        hipStream_t st = mstream(M);
        if (!hl.empty()) {
            HIPCHK(hipMemcpyAsync(M->hub_list, hl.data(), hl.size() * 4, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(M->hub_items, hi.data(), hi.size() * 4, hipMemcpyHostToDevice, st));
        }
        HIPCHK(hipStreamSynchronize(st));
        M->n_hub = (uint32_t)hl.size();
        M->n_hub_items = (uint32_t)(hi.size() / 2);
        M->hub_uid = G->uid;
    }
    gst_status sst = fwd_bufs(M, M->f[0], n);
    if (sst != GST_OK) return sst;
    if (M->static_uid != G->uid) {
        // graph-only layer-1 means: the gather over features with every group bit clear (set 0's
        // buffers as scratch, on its stream; the other set's stream waits for e_static)
        FwdSet& F = M->f[0];
        hipStream_t st = fstream(M, F);
        const int fg = std::max(1, ctx->cu_count * 8);
        k_fill32<<<gridn(n, 256, fg), 256, 0, st>>>(F.gm, 0u, (uint64_t)n);
        k_features<<<gridn(n, 256, fg), 256, 0, st>>>(G->d_row, G->d_w, G->w_type == GST_W_FLOAT32 ? 1 : 0, F.gm, n, F.x);
        aggregate<FIN, false>(M, G, F.hub_acc, F.x, M->ax_static, st);
        if (!M->e_static) HIPCHK(hipEventCreateWithFlags(&M->e_static, hipEventDisableTiming));
        HIPCHK(hipEventRecord(M->e_static, st));
        HIPCHK(hipGetLastError());
        M->static_uid = G->uid;
    }
    return GST_OK;
}

// pinned (may be nullptr): a staging buffer of at least g + 1 + members words that stays untouched
// until the stream has run the copy; without it the copy is waited for
gst_status upload_groups(gst_model* M, FwdSet& F, const gst_graph* G, const gst_groups* groups, int32_t* pinned = nullptr) {
    gst_ctx* ctx = M->ctx;
    hipStream_t st = fstream(M, F);
    const int g = groups->n_groups;
    const uint32_t cnt = (uint32_t)(g + 1 + groups->grp_ptr[g]);
    if (F.cap_grp < cnt) {
        HIPCHK(hipStreamSynchronize(st));   // earlier selections may still read the old buffer
        if (!dalloc(F.grp, cnt)) { F.cap_grp = 0; return gst_set_error(ctx, GST_OOM, "model: group allocation failed"); }
        F.cap_grp = cnt;
    }
    std::vector<int32_t> hv;
    int32_t* h = pinned;
    if (!h) { hv.resize(cnt); h = hv.data(); }
    for (int i = 0; i <= g; ++i) h[(size_t)i] = groups->grp_ptr[i];
    for (int32_t j = 0; j < groups->grp_ptr[g]; ++j) h[(size_t)(g + 1 + j)] = groups->grp_vtx[j];
    HIPCHK(hipMemcpyAsync(F.grp, h, cnt * 4, hipMemcpyHostToDevice, st));
    if (!pinned) HIPCHK(hipStreamSynchronize(st));   // h is a local buffer
    const int fg = std::max(1, ctx->cu_count * 8);
    k_fill32<<<gridn(G->n, 256, fg), 256, 0, st>>>(F.gm, 0u, (uint64_t)G->n);
    k_set_bits_scatter<<<gridn((uint64_t)groups->grp_ptr[g] * 64, 256, fg), 256, 0, st>>>(
        F.gm, F.grp, F.grp + g + 1, g, G->d_row, G->d_col, F.gcnt);   // features use bits 0..7
    return GST_OK;
}

template <int D, bool BWD>
void aggregate(gst_model* M, const gst_graph* G, float* hub_acc, const void* in, void* out, hipStream_t st) {
    const int fg = std::max(1, M->ctx->cu_count * 8);
    if (M->n_hub) k_fill32<<<gridn((uint64_t)M->n_hub * D, 256, fg), 256, 0, st>>>((uint32_t*)hub_acc, 0u,
                                                                               (uint64_t)M->n_hub * D);
    if (BWD)
        k_aggregate_bwd<D, BWD><<<std::max(1, M->ctx->cu_count * 16), 256, 0, st>>>(G->d_row, G->d_col, (uint32_t)G->n, in,
                                                                                  out, hub_acc, M->hub_items, M->n_hub_items);
    else
        k_aggregate_fwd<D><<<std::max(1, M->ctx->cu_count * 16), 256, 0, st>>>(G->d_row, G->d_col, (uint32_t)G->n,
                                                                             (const bf16*)in, (bf16*)out, hub_acc,
                                                                             M->hub_items, M->n_hub_items);
    if (M->n_hub)
        k_aggregate_hub_finish<D, BWD><<<gridn((uint64_t)M->n_hub * D, 256, fg), 256, 0, st>>>(G->d_row, M->hub_list,
                                                                                        M->n_hub, hub_acc, out);
}

// the forward pass of one query into activation set F, on F's stream.  need_h2: keep the layer-2
// activations (the backward pass reads them; set 0 only); the scores alone need not
gst_status forward(gst_model* M, FwdSet& F, const gst_graph* G, const gst_groups* groups,
                   int32_t* pinned_groups = nullptr, bool need_h2 = false) {
    gst_ctx* ctx = M->ctx;
    gst_status sst = model_bufs(M, G);
    if (sst != GST_OK) return sst;
    sst = fwd_bufs(M, F, (uint32_t)G->n);
    if (sst != GST_OK) return sst;
    hipStream_t st = fstream(M, F);
    if (st != fstream(M, M->f[0])) HIPCHK(hipStreamWaitEvent(st, M->e_static, 0));
    sst = upload_groups(M, F, G, groups, pinned_groups);
    if (sst != GST_OK) return sst;
    const uint32_t n = (uint32_t)G->n;
    const int fg = std::max(1, ctx->cu_count * 8);
    const int dg = std::max(1, ctx->cu_count * GST_DENSE_GRID);
    k_features<<<gridn(n, 256, fg), 256, 0, st>>>(G->d_row, G->d_w, G->w_type == GST_W_FLOAT32 ? 1 : 0, F.gm, n, F.x);
    k_ax_finish<<<gridn(n, 256, fg), 256, 0, st>>>(G->d_row, M->ax_static, F.gcnt, n, F.ax);
    k_dense_fwd<FIN, false><<<dg, 256, 0, st>>>(F.x, F.ax, M->W1t, M->p + OFF_B1, n, F.h1, nullptr, nullptr, nullptr);
    aggregate<HID, false>(M, G, F.hub_acc, F.h1, F.ah1, st);
    k_dense_fwd<HID, true><<<dg, 256, 0, st>>>(F.h1, F.ah1, M->W2t, M->p + OFF_B2, n, need_h2 ? M->h2 : nullptr,
                                               M->p + OFF_W3, M->p + OFF_B3, F.s);
    HIPCHK(hipGetLastError());
    return GST_OK;
}

}  // namespace

// ---------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------
extern "C" {

int64_t gst_model_num_params(void) { return NPARAM; }

gst_status gst_model_create(gst_ctx* ctx, uint64_t seed, gst_model** out) {
    if (!ctx || !out) return GST_INVALID_ARG;
    *out = nullptr;
    if (!ctx->stream) return gst_set_error(ctx, GST_NO_DEVICE, "context has no device");
    HIPCHK(hipSetDevice(ctx->device));
    gst_model* M = new gst_model();
    M->ctx = ctx;
    bool ok = dalloc(M->p, NPARAM) && dalloc(M->g, NPARAM) && dalloc(M->m, NPARAM) && dalloc(M->v, NPARAM) &&
              dalloc(M->W1t, HID * K1) && dalloc(M->W2t, HID * K2) && dalloc(M->W1kn, HID * K1) &&
              dalloc(M->W2kn, HID * K2) && dalloc(M->loss, 1);
    if (!ok) { gst_model_free(M); return gst_set_error(ctx, GST_OOM, "model: parameter allocation failed"); }
    // He-style init from the library's own generator (splitmix64 -> uniform)
    std::vector<float> h(NPARAM, 0.f);
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
    auto unif = [&]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    for (int i = 0; i < HID * K1; ++i) h[OFF_W1 + i] = (float)(unif() * std::sqrt(6.0 / 20.0));
    for (int i = 0; i < HID * K2; ++i) h[OFF_W2 + i] = (float)(unif() * std::sqrt(6.0 / K2));
    for (int i = 0; i < HID; ++i) h[OFF_W3 + i] = (float)(unif() * std::sqrt(6.0 / HID));
    gst_status s = gst_model_set_params(M, h.data());
    if (s != GST_OK) { gst_model_free(M); return s; }
    *out = M;
    return GST_OK;
}

void gst_model_free(gst_model* M) {
    if (!M) return;
    hipFree(M->p); hipFree(M->g); hipFree(M->m); hipFree(M->v);
    hipFree(M->W1t); hipFree(M->W2t); hipFree(M->W1kn); hipFree(M->W2kn); hipFree(M->loss);
    hipFree(M->ax_static);
    for (FwdSet& F : M->f) {
        hipFree(F.x); hipFree(F.ax); hipFree(F.h1); hipFree(F.ah1); hipFree(F.s); hipFree(F.gm); hipFree(F.gcnt);
        hipFree(F.hub_acc); hipFree(F.grp);
    }
    if (M->e_static) hipEventDestroy(M->e_static);
    hipFree(M->h2); hipFree(M->dZ2); hipFree(M->dZ1);
    hipFree(M->dS); hipFree(M->dX2); hipFree(M->dH1); hipFree(M->dA1);
    hipFree(M->y); hipFree(M->hub_list); hipFree(M->hub_items);
    sel_free(M->sync_sel);
    for (auto& sl : M->batch_sel) sel_free(sl);
    for (auto& b : M->sub_pool) sub_free(b);
    sub_free(M->sync_buf);
    delete M;
}

gst_status gst_model_set_params(gst_model* M, const float* in) {
    if (!M || !in) return GST_INVALID_ARG;
    gst_ctx* ctx = M->ctx;
    hipStream_t st = (hipStream_t)ctx->stream;
    HIPCHK(hipMemcpyAsync(M->p, in, NPARAM * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemsetAsync(M->g, 0, NPARAM * 4, st));
    HIPCHK(hipMemsetAsync(M->m, 0, NPARAM * 4, st));
    HIPCHK(hipMemsetAsync(M->v, 0, NPARAM * 4, st));
    k_refresh_bf16<<<64, 256, 0, st>>>(M->p, M->W1t, M->W2t, M->W1kn, M->W2kn);
    HIPCHK(hipStreamSynchronize(st));
    M->steps = 0;
    M->grad_count = 0;
    return GST_OK;
}

gst_status gst_model_get_params(gst_model* M, float* out) {
    if (!M || !out) return GST_INVALID_ARG;
    gst_ctx* ctx = M->ctx;
    hipStream_t st = (hipStream_t)ctx->stream;
    HIPCHK(hipMemcpyAsync(out, M->p, NPARAM * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return GST_OK;
}

gst_status gst_model_scores(gst_model* M, const gst_graph* G, const gst_groups* groups, float* scores) {
    if (!M || !G || !groups || !scores) return GST_INVALID_ARG;
    gst_ctx* ctx = M->ctx;
    gst_status s = forward(M, M->f[0], G, groups);
    if (s != GST_OK) return s;
    hipStream_t st = (hipStream_t)ctx->stream;
    HIPCHK(hipMemcpyAsync(scores, M->f[0].s, (size_t)G->n * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return GST_OK;
}

gst_status gst_model_loss_grad(gst_model* M, const gst_graph* G, const gst_groups* groups, const uint8_t* labels,
                               float pos_weight, float* loss_out, float* grad_out) {
    if (!M || !G || !groups || !labels) return GST_INVALID_ARG;
    gst_ctx* ctx = M->ctx;
    gst_status sst = forward(M, M->f[0], G, groups, nullptr, true);
    if (sst != GST_OK) return sst;
    hipStream_t st = (hipStream_t)ctx->stream;
    const FwdSet& F = M->f[0];
    const uint32_t n = (uint32_t)G->n;
    const int fg = std::max(1, ctx->cu_count * 8);
    const int dg = std::max(1, ctx->cu_count * 8);
    HIPCHK(hipMemcpyAsync(M->y, labels, n, hipMemcpyHostToDevice, st));
    k_fill32<<<1, 64, 0, st>>>((uint32_t*)M->loss, 0u, 1);
    if (grad_out) k_fill32<<<gridn(NPARAM, 256, fg), 256, 0, st>>>((uint32_t*)M->g, 0u, NPARAM);
    k_loss<<<gridn(n, 256, fg), 256, 0, st>>>(F.s, M->y, n, pos_weight, M->dS, M->loss);
    k_head_bwd<<<gridn((uint64_t)n * HID, 256, fg), 256, 0, st>>>(M->dS, M->h2, M->p + OFF_W3, n, M->dZ2,
                                                                   M->g + OFF_W3, M->g + OFF_B3);
    k_bias_grad<<<gridn((uint64_t)n * HID, 256, fg), 256, 0, st>>>(M->dZ2, n, M->g + OFF_B2);
    k_dense_bwd_w<HID><<<std::max(1, ctx->cu_count * 2), 256, 0, st>>>(M->dZ2, F.h1, F.ah1, n, M->g + OFF_W2);
    k_dense_bwd_x<K2><<<dg, 256, 0, st>>>(M->dZ2, M->W2kn, n, M->dX2);
    k_split_self<<<gridn((uint64_t)n * HID, 256, fg), 256, 0, st>>>(M->dX2, n, M->dH1, M->dA1);
    aggregate<HID, true>(M, G, F.hub_acc, M->dA1, M->dH1, st);
    k_relu_mask_bias<<<gridn((uint64_t)n * HID, 256, fg), 256, 0, st>>>(M->dH1, F.h1, n, M->dZ1, M->g + OFF_B1);
    k_dense_bwd_w<FIN><<<std::max(1, ctx->cu_count * 2), 256, 0, st>>>(M->dZ1, F.x, F.ax, n, M->g + OFF_W1);
    HIPCHK(hipGetLastError());
    float l = 0;
    HIPCHK(hipMemcpyAsync(&l, M->loss, 4, hipMemcpyDeviceToHost, st));
    if (grad_out) HIPCHK(hipMemcpyAsync(grad_out, M->g, NPARAM * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (loss_out) *loss_out = l;
    M->grad_count++;
    return GST_OK;
}

gst_status gst_model_adam_step(gst_model* M, float lr) {
    if (!M) return GST_INVALID_ARG;
    gst_ctx* ctx = M->ctx;
    if (M->grad_count == 0) return GST_OK;
    hipStream_t st = (hipStream_t)ctx->stream;
    M->steps++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.f - std::pow(b1, (float)M->steps), bc2 = 1.f - std::pow(b2, (float)M->steps);
    k_adam<<<gridn(NPARAM, 256, 64), 256, 0, st>>>(M->p, M->g, M->m, M->v, NPARAM, lr, b1, b2, 1e-8f, bc1, bc2,
                                                   1.f / (float)M->grad_count, M->W1t, M->W2t, M->W1kn, M->W2kn);
    k_refresh_bf16<<<64, 256, 0, st>>>(M->p, M->W1t, M->W2t, M->W1kn, M->W2kn);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    M->grad_count = 0;
    return GST_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// learned-bound solve
// ---------------------------------------------------------------------------------------
namespace {

// Scores, top-k selection and the subgraph induced by the top-k vertices plus every group member,
// for one query, in two parts on the model's stream.  select_begin queues the network, the top-k
// and the subgraph's sizes and group lists (readback into sl.h_rb, event sl.e_sizes); nothing waits.
// select_finish (once e_sizes is complete) reserves buf for those sizes, queues the subgraph's CSR
// (event sl.e_built) and fills S as a view of buf (never freed as a graph) and sub_vtx with the
// members' subgraph ids.  A slot's scratch is reused only after its e_built.
gst_status select_begin(gst_model* M, FwdSet& F, const gst_graph* G, const gst_groups* groups, double frac,
                        SelSlot& sl) {
    gst_ctx* ctx = M->ctx;
    hipStream_t st = fstream(M, F);
    sl.st = st;
    const uint32_t n = (uint32_t)G->n;
    const int fg = std::max(1, ctx->cu_count * 8);
    const int g = groups->n_groups;
    const uint32_t gcnt = (uint32_t)groups->grp_ptr[g];
    const uint32_t nbk = (n + 1023) / 1024;
    const uint64_t need = 2048ull + 8 + 5ull * n + 2ull * (nbk + 2) + 2 + gcnt + 64;
    if (sl.cap < need) {
        if (!dalloc(sl.scr, need)) { sl.cap = 0; return gst_set_error(ctx, GST_OOM, "learned solve: scratch allocation failed"); }
        sl.cap = need;
        k_fill32<<<8, 256, 0, st>>>(sl.scr, 0u, 2048);   // the histogram starts clear; each pass clears it
    }
    if (sl.h_cap < gcnt + 2) {
        if (sl.h_rb) hipHostFree(sl.h_rb);
        sl.h_rb = nullptr;
        sl.h_cap = 0;
        if (hipHostMalloc((void**)&sl.h_rb, (gcnt + 2 + 1024) * 4ull) != hipSuccess)
            return gst_set_error(ctx, GST_OOM, "learned solve: pinned allocation failed");
        sl.h_cap = gcnt + 2 + 1024;
    }
    if (sl.hg_cap < (uint32_t)g + 1 + gcnt) {
        if (sl.h_grp) hipHostFree(sl.h_grp);
        sl.h_grp = nullptr;
        sl.hg_cap = 0;
        if (hipHostMalloc((void**)&sl.h_grp, ((size_t)g + 1 + gcnt + 1024) * 4ull) != hipSuccess)
            return gst_set_error(ctx, GST_OOM, "learned solve: pinned allocation failed");
        sl.hg_cap = (uint32_t)g + 1 + gcnt + 1024;
    }
    if (!sl.e_sizes) HIPCHK(hipEventCreateWithFlags(&sl.e_sizes, hipEventDisableTiming));
    if (!sl.e_built) HIPCHK(hipEventCreateWithFlags(&sl.e_built, hipEventDisableTiming));
    sl.n = n; sl.nbk = nbk; sl.gcnt = gcnt;
    gst_status sst = forward(M, F, G, groups, sl.h_grp);
    if (sst != GST_OK) return sst;
    uint32_t k_target = (uint32_t)std::min<double>(n, std::ceil(frac * n));
    if (k_target < 1) k_target = 1;
    uint32_t* d_hist = sl.scr;
    uint32_t* d_state = d_hist + 2048;
    uint32_t* d_sel = d_state + 8;
    uint32_t* d_newid = d_sel + n;
    uint32_t* d_deg = d_newid + n;
    uint32_t* d_arcoff = d_deg + n;
    uint32_t* d_old = d_arcoff + n;
    uint32_t* d_bsum = d_old + n;        // 2 * (nbk + 2)
    uint32_t* d_hdr = d_bsum + 2 * (nbk + 2);   // vertex count, arc count, then the members' new ids
    int32_t* d_subgrp = (int32_t*)(d_hdr + 2);
    // ---- top-k threshold on the device (the first pass starts the state) ----
    const int hg = std::max(1, std::min(ctx->cu_count, (int)((n + 4095) / 4096)));
    for (int pass = 0; pass < 3; ++pass) {
        k_key_hist<<<hg, 256, 0, st>>>(F.s, n, pass, d_state, d_hist);
        k_pick_bin<<<1, 256, 0, st>>>(d_hist, d_state, pass, k_target);
    }
    // ---- induced subgraph: selection and new ids, arc offsets, group lists in new ids ----
    k_select_scan<<<nbk, 256, 0, st>>>(F.s, F.gm, n, d_state, d_sel, d_newid, d_bsum);
    k_scan_sums<<<1, 64, 0, st>>>(d_bsum, nbk, d_hdr);
    k_scan_add_list<<<gridn(n, 256, fg), 256, 0, st>>>(d_newid, n, d_bsum, d_sel, d_old, d_deg);
    k_sub_deg<<<std::max(1, ctx->cu_count * 8), 256, 0, st>>>(G->d_row, G->d_col, d_sel, d_old, d_bsum + nbk, d_deg,
                                                              F.grp + g + 1, gcnt, d_newid, d_subgrp);
    k_scan_blocks<<<nbk, 256, 0, st>>>(d_deg, n, d_arcoff, d_bsum + nbk + 2);
    k_scan_sums<<<1, 64, 0, st>>>(d_bsum + nbk + 2, nbk, d_hdr + 1);   // offsets completed by k_sub_fill
    HIPCHK(hipMemcpyAsync(sl.h_rb, d_hdr, (2ull + gcnt) * 4ull, hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(sl.e_sizes, st));
    HIPCHK(hipGetLastError());
    return GST_OK;
}

gst_status select_finish(gst_model* M, const gst_graph* G, SelSlot& sl, SubBuf& buf, gst_graph& S,
                         std::vector<int32_t>& sub_vtx, gst_learned_stats& L, bool pooled = false) {
    gst_ctx* ctx = M->ctx;
    hipStream_t st = sl.st;   // the stream its select_begin ran on
    const uint32_t n = sl.n, nbk = sl.nbk, gcnt = sl.gcnt;
    const uint32_t ns = sl.h_rb[0], ms = sl.h_rb[1];
    sub_vtx.assign((const int32_t*)&sl.h_rb[2], (const int32_t*)&sl.h_rb[2] + gcnt);
    L.sub_vertices = ns;
    L.sub_arcs = ms;
    if (!sub_reserve(buf, ns, ms, pooled ? M->sub_max_n : 0u, pooled ? M->sub_max_m : 0ull))
        return gst_set_error(ctx, GST_OOM, "learned solve: subgraph allocation failed");
    if (pooled) {
        M->sub_max_n = std::max(M->sub_max_n, buf.cap_n);
        M->sub_max_m = std::max(M->sub_max_m, buf.cap_m);
    }
    S.ctx = ctx;
    S.uid = gst_next_graph_uid();
    S.budget_key = G->uid | (1ull << 63);   // learned subgraphs of one graph share round budgets
    S.n = (int32_t)ns;
    S.m = ms;
    S.w_type = G->w_type;
    // the solver reads rows and arc records only (the network never runs on a subgraph)
    S.d_row = buf.row; S.d_arc = buf.arc; S.d_col = nullptr; S.d_w = nullptr;
    S.max_deg = 0;
    if (ns > 0) {
        uint32_t* d_sel = sl.scr + 2048 + 8;
        uint32_t* d_newid = d_sel + n;
        uint32_t* d_arcoff = d_newid + 2 * (size_t)n;
        uint32_t* d_old = d_arcoff + n;
        uint32_t* d_bsum = d_old + n;
        k_sub_fill<<<gridn((uint64_t)ns * 64, 256, ctx->cu_count * 8), 256, 0, st>>>(
            G->d_row, G->d_arc, d_sel, d_newid, d_arcoff, d_bsum + nbk + 2, d_bsum + nbk + 2 + nbk, d_old, ns,
            S.d_row, S.d_arc);
    }
    HIPCHK(hipEventRecord(sl.e_built, st));
    HIPCHK(hipGetLastError());
    return GST_OK;
}

// both parts, waited for (the one-query path); with want_host_rows the host row offsets are
// copied too (the one-query solver reads them)
gst_status learned_subgraph(gst_model* M, const gst_graph* G, const gst_groups* groups, double frac, SubBuf& buf,
                            gst_graph& S, std::vector<int32_t>& sub_vtx, gst_learned_stats& L, bool want_host_rows) {
    S.n = 0;   // no subgraph unless one is built below
    if (sub_declined(M, G)) return GST_OK;
    gst_status sst = select_begin(M, M->f[0], G, groups, frac, M->sync_sel);
    if (sst != GST_OK) return sst;
    gst_ctx* ctx = M->ctx;
    HIPCHK(hipEventSynchronize(M->sync_sel.e_sizes));
    if (sub_oversized(M, G, M->sync_sel.h_rb[1])) {
        L.sub_vertices = M->sync_sel.h_rb[0];
        L.sub_arcs = M->sync_sel.h_rb[1];
        return GST_OK;
    }
    sst = select_finish(M, G, M->sync_sel, buf, S, sub_vtx, L);
    if (sst != GST_OK) return sst;
    HIPCHK(hipEventSynchronize(M->sync_sel.e_built));
    if (want_host_rows && S.n > 0) {
        std::vector<uint32_t> r32((size_t)S.n + 1);
        HIPCHK(hipMemcpyAsync(r32.data(), S.d_row, ((size_t)S.n + 1) * 4ull, hipMemcpyDeviceToHost, mstream(M)));
        HIPCHK(hipStreamSynchronize(mstream(M)));
        S.h_row.assign(r32.begin(), r32.end());
    }
    return GST_OK;
}

}  // namespace

gst_status gst_learned_solve(gst_ctx* ctx, const gst_graph* G, const gst_groups* groups, gst_model* M, double frac,
                             const gst_opts* opts_in, gst_result* res, gst_learned_stats* ls) {
    hipStream_t st = (hipStream_t)ctx->stream;
    gst_learned_stats L{};
    struct Events {   // destroyed on every return
        hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
        ~Events() { for (auto x : e) if (x) hipEventDestroy(x); }
    } ev;
    for (auto& x : ev.e) HIPCHK(hipEventCreate(&x));
    const hipEvent_t e0 = ev.e[0], e1 = ev.e[1], e2 = ev.e[2], e3 = ev.e[3];
    HIPCHK(hipEventRecord(e0, st));
    SubBuf& buf = M->sync_buf;   // reused while large enough (an allocation costs more than a selection)
    gst_graph S;
    std::vector<int32_t> sub_vtx;
    gst_status sst = learned_subgraph(M, G, groups, frac, buf, S, sub_vtx, L, true);
    if (sst != GST_OK) return sst;
    const uint32_t ns = (uint32_t)S.n;
    HIPCHK(hipEventRecord(e1, st));
    const int g = groups->n_groups;
    gst_groups sg{g, groups->grp_ptr, sub_vtx.data()};
    gst_opts so{};
    so.prune_lb = 1;
    so.want_tree = 0;   // the bound is a cost
    so.auto_bound = 1;
    gst_result sr;
    std::memset(&sr, 0, sizeof(sr));
    double ub = 0.0;
    bool have_ub = false;
    if (ns > 0) {
        gst_status ss = gst_dpbf_solve(ctx, &S, &sg, &so, &sr);
        if (ss == GST_OK) { ub = sr.cost; have_ub = true; }
        gst_result_free(&sr);
    }
    HIPCHK(hipEventRecord(e2, st));
    HIPCHK(hipStreamSynchronize(st));
    // ---- full solve, bounded by the subgraph tree when there is one ----
    gst_opts fo{};
    if (opts_in) fo = *opts_in;
    else { fo.prune_lb = 1; fo.want_tree = 1; }
    if (have_ub) fo.ub = ub;
    fo.auto_bound = 1;   // the spider bound tightens the subgraph bound where it is lower (as in the batch)
    gst_status fs = gst_dpbf_solve(ctx, G, groups, &fo, res);
    if (fs == GST_INFEASIBLE && have_ub) {
        // the bound cannot be below the optimum; reaching here means float rounding in the bound:
        // solve without it
        fo.ub = 0;
        fs = gst_dpbf_solve(ctx, G, groups, &fo, res);
        if (fs == GST_OK) res->fell_back = 1;
    } else if (fs == GST_OK && !have_ub) {
        res->fell_back = 1;
    }
    HIPCHK(hipEventRecord(e3, st));
    HIPCHK(hipEventSynchronize(e3));
    float a = 0, b = 0, c = 0;
    hipEventElapsedTime(&a, e0, e1);
    hipEventElapsedTime(&b, e1, e2);
    hipEventElapsedTime(&c, e2, e3);
    L.t_score_select_ms = a;
    L.t_sub_solve_ms = b;
    L.t_full_solve_ms = c;
    L.ub = have_ub ? ub : -1.0;
    if (ls) *ls = L;
    return fs;
}

// ---------------------------------------------------------------------------------------
// batched learned-bound solve: per query a subgraph job and then a bounded full job, both run by
// the batched engine, so the subgraph solves and full solves of many queries share launches
// ---------------------------------------------------------------------------------------
gst_status gst_learned_solve_batch(gst_ctx* ctx, const gst_graph* G, int32_t nq, const gst_groups* groups,
                                   gst_model* M, double frac, const gst_opts* opts, const gst_batch_opts* bopts,
                                   gst_result* results, gst_status* status, gst_learned_stats* lstats,
                                   gst_batch_stats* bstats) {
    struct Q {
        gst_graph S;
        int buf = -1;                   // subgraph buffer in use (index into bufs)
        std::vector<int32_t> sub_vtx;
        gst_groups sg{};
        gst_result sub_res{};
        gst_learned_stats L{};
        double ub = 0.0;
        double t_sel = 0.0;             // host ms when its selection was queued
        bool have_ub = false, retried = false;
    };
    std::vector<Q> qs((size_t)std::max(nq, 0));
    // subgraph CSRs: the model's pool, kept across calls (an allocation or free costs far more than a
    // selection; a pool buffer is in use from its selection's build to its subgraph job's end)
    std::vector<SubBuf>& bufs = M->sub_pool;
    std::vector<int> free_bufs;
    for (int i = (int)bufs.size() - 1; i >= 0; --i) free_bufs.push_back(i);
    std::vector<int32_t> pending_full;   // queries whose full solve may start (engine thread only)
    int gmax = 1;
    for (int32_t q = 0; q < nq; ++q)
        if (status[q] == GST_OK) gmax = std::max(gmax, groups[q].n_groups);
    gst_status setup_err = GST_OK;
    // the selection of upcoming queries runs on streams of its own while the engine's step runs: set 0
    // of the activations on stream2, set 1 on stream3, so that two selections overlap (each is some
    // 20 short launches whose gaps the other fills)
    static_assert(GST_SELECT_LANES >= 1 && GST_SELECT_LANES <= FWD_SETS, "selection lanes");
    const int lanes = GST_SELECT_LANES;
    if (!ctx->stream2) {
        hipStream_t s2;
        HIPCHK(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
        ctx->stream2 = (void*)s2;
    }
    if (lanes > 1 && !ctx->stream3) {
        hipStream_t s3;
        HIPCHK(hipStreamCreateWithFlags(&s3, hipStreamNonBlocking));
        ctx->stream3 = (void*)s3;
    }
    const int lookahead = GST_SELECT_AHEAD;
    M->sel_stream = ctx->stream2;   // set 0's selections (and the model's set-up) queue on stream2
    if (lanes > 1) M->f[1].st = (hipStream_t)ctx->stream3;
    auto end_streams = [&]() {
        M->sel_stream = nullptr;
        M->f[1].st = nullptr;
    };
    auto now = []() {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
    };
    auto full_job = [&](int32_t q, gst_batch_job& j) {
        Q& x = qs[(size_t)q];
        j = gst_batch_job();
        j.G = G;
        j.groups = &groups[q];
        if (opts) j.opts = opts[q];
        else { j.opts = gst_opts(); j.opts.prune_lb = 1; j.opts.want_tree = 1; }
        j.opts.ub = x.have_ub && !x.retried ? x.ub : 0.0;
        j.opts.auto_bound = 1;   // the spider bound tightens the subgraph bound where it is lower
        j.res = &results[q];
        j.tag = q;
        j.kind = 1;
    };
    // Selection runs on a host thread of its own: a query's network pass, top-k and subgraph build
    // are ~30 launches on the selection stream, and the engine's thread only fills slots and steps.
    // Under mu: the queues of selected queries, the free subgraph buffers, the count of selections
    // begun and not yet taken by the engine (at most `lookahead`), and the selection thread's end.
    // selections queued at once (the oldest finishes first); slot i runs on activation set i % lanes
    const size_t depth = std::max<size_t>(GST_SELECT_DEPTH, (size_t)lanes);
    if (M->batch_sel.size() < depth) M->batch_sel.resize(depth);
    std::mutex mu;
    std::condition_variable cv;
    std::deque<int32_t> ready_q;   // subgraph built: its subgraph job may start
    std::deque<int32_t> nosub_q;   // no subgraph (empty, or no memory for it): the plain solve runs
    size_t untaken = 0;
    bool sel_end = false, stop = false;
    auto selector = [&]() {
        gst_status err = GST_OK;
        struct Sel { SelSlot* sl; int32_t q; };
        std::deque<Sel> inflight;
        std::vector<SelSlot*> idle_slots;
        for (size_t i = 0; i < depth; ++i) idle_slots.push_back(&M->batch_sel[i]);
        auto publish = [&](std::deque<int32_t>& to, int32_t q) {
            { std::lock_guard<std::mutex> lk(mu); to.push_back(q); }
            cv.notify_all();
        };
        // the oldest selection to its end: sizes, subgraph build, then ready
        auto finish_oldest = [&]() -> gst_status {
            const Sel e = inflight.front();
            inflight.pop_front();
            Q& x = qs[(size_t)e.q];
            HIPCHK(hipEventSynchronize(e.sl->e_sizes));
            if (sub_oversized(M, G, e.sl->h_rb[1])) {
                // too large to pay for itself: no subgraph, the spider-bounded full solve runs
                x.L.sub_vertices = e.sl->h_rb[0];
                x.L.sub_arcs = e.sl->h_rb[1];
                idle_slots.push_back(e.sl);
                publish(nosub_q, e.q);
                return GST_OK;
            }
            int bi = -1;
            {
                std::lock_guard<std::mutex> lk(mu);
                if (!free_bufs.empty()) { bi = free_bufs.back(); free_bufs.pop_back(); }
            }
            if (bi < 0) { bufs.emplace_back(); bi = (int)bufs.size() - 1; }   // only this thread grows the pool
            x.buf = bi;
            const gst_status ss = select_finish(M, G, *e.sl, bufs[(size_t)bi], x.S, x.sub_vtx, x.L, true);
            // the slot's scratch is free once its (possibly empty) build has run
            HIPCHK(hipEventSynchronize(e.sl->e_built));
            idle_slots.push_back(e.sl);
            if (ss != GST_OK || x.S.n == 0) {
                {
                    std::lock_guard<std::mutex> lk(mu);
                    free_bufs.push_back(bi);
                }
                x.buf = -1;
                if (ss != GST_OK && ss != GST_OOM) { status[e.q] = ss; return ss; }
                publish(nosub_q, e.q);
                return GST_OK;
            }
            x.L.t_score_select_ms = now() - x.t_sel;
            x.sg = gst_groups{groups[e.q].n_groups, groups[e.q].grp_ptr, x.sub_vtx.data()};
            publish(ready_q, e.q);
            return GST_OK;
        };
        int last_lane = -1;
        auto run = [&]() -> gst_status {
            HIPCHK(hipSetDevice(ctx->device));
            int32_t next = 0;   // queries below next were begun or failed validation
            for (;;) {
                while (next < nq && status[next] != GST_OK) ++next;
                bool begin = false;
                {
                    std::unique_lock<std::mutex> lk(mu);
                    if (stop) return GST_OK;
                    begin = next < nq && !idle_slots.empty() && untaken < (size_t)lookahead;
                    if (begin) ++untaken;
                    else if (inflight.empty()) {
                        if (next >= nq) return GST_OK;
                        // the engine holds `lookahead` selected queries: wait until it takes one
                        cv.wait(lk, [&] { return stop || untaken < (size_t)lookahead; });
                        continue;
                    }
                }
                if (begin) {
                    const int32_t q = next++;
                    if (sub_declined(M, G)) {   // the model no longer selects on this graph
                        publish(nosub_q, q);
                        continue;
                    }
                    qs[(size_t)q].t_sel = now();
                    // an idle slot on another set than the last selection's, so that the sets alternate
                    size_t pick = idle_slots.size() - 1;
                    for (size_t i = 0; i < idle_slots.size(); ++i)
                        if ((int)((idle_slots[i] - M->batch_sel.data()) % lanes) != last_lane) { pick = i; break; }
                    SelSlot* sl = idle_slots[pick];
                    idle_slots.erase(idle_slots.begin() + (ptrdiff_t)pick);
                    last_lane = (int)((sl - M->batch_sel.data()) % lanes);
                    const gst_status ss = select_begin(M, M->f[last_lane], G, &groups[q], frac, *sl);
                    if (ss == GST_OK) { inflight.push_back({sl, q}); continue; }
                    idle_slots.push_back(sl);
                    if (ss != GST_OOM) { status[q] = ss; return ss; }
                    publish(nosub_q, q);   // no subgraph: the plain solve runs (a fallback)
                    continue;
                }
                const gst_status fs = finish_oldest();
                if (fs != GST_OK) return fs;
            }
        };
        err = run();
        {
            std::lock_guard<std::mutex> lk(mu);
            if (err != GST_OK && setup_err == GST_OK) setup_err = err;
            sel_end = true;
        }
        cv.notify_all();
    };
    auto next_job = [&](gst_batch_job& j, bool wait) -> bool {
        if (!pending_full.empty()) {   // the engine's thread alone adds and takes these
            const int32_t q = pending_full.back();
            pending_full.pop_back();
            full_job(q, j);
            return true;
        }
        int32_t q = -1;
        bool sub = false;
        {
            std::unique_lock<std::mutex> lk(mu);
            // nothing ready: the engine steps on with the queries it holds, or (no slot busy) waits
            if (wait) cv.wait(lk, [&] { return !ready_q.empty() || !nosub_q.empty() || sel_end; });
            if (!nosub_q.empty()) { q = nosub_q.front(); nosub_q.pop_front(); }
            else if (!ready_q.empty()) { q = ready_q.front(); ready_q.pop_front(); sub = true; }
            else return false;
            --untaken;
        }
        cv.notify_all();
        if (!sub) { full_job(q, j); return true; }
        Q& x = qs[(size_t)q];
        j = gst_batch_job();
        j.G = &x.S;
        j.groups = &x.sg;
        j.opts = gst_opts();
        j.opts.prune_lb = 1;
        j.opts.want_tree = 0;
        j.opts.auto_bound = 1;
        j.temp_graph = true;
        j.res = &x.sub_res;
        j.tag = q;
        j.kind = 0;
        return true;
    };
    std::vector<uint8_t> done((size_t)std::max(nq, 0), 0);   // the query's full solve finished
    auto job_done = [&](gst_batch_job& j) {
        const int32_t q = (int32_t)j.tag;
        Q& x = qs[(size_t)q];
        if (j.kind == 0) {
            x.L.t_sub_solve_ms = j.t_done - j.t_fill;
            if (j.status == GST_OK) { x.have_ub = true; x.ub = x.sub_res.cost; }
            gst_result_free(&x.sub_res);
            if (x.buf >= 0) {
                std::lock_guard<std::mutex> lk(mu);
                free_bufs.push_back(x.buf);
            }
            x.buf = -1;
            pending_full.push_back(q);
            return;
        }
        x.L.t_full_solve_ms += j.t_done - j.t_fill;
        if (j.status == GST_INFEASIBLE && x.have_ub && !x.retried) {
            // the bound cannot be below the optimum; only float rounding gets here: run unbounded
            x.retried = true;
            gst_result_free(&results[q]);
            pending_full.push_back(q);
            return;
        }
        status[q] = j.status;
        done[(size_t)q] = 1;
        if (j.status == GST_OK) results[q].fell_back = (!x.have_ub || x.retried) ? 1 : 0;
        x.L.ub = x.have_ub ? x.ub : -1.0;
        if (lstats) lstats[q] = x.L;
    };
    int64_t valid = 0;   // a query has one job (its subgraph's or its full solve) in flight at a time
    for (int32_t q = 0; q < nq; ++q) valid += status[q] == GST_OK ? 1 : 0;
    std::thread sel_th;
    try {
        sel_th = std::thread(selector);
    } catch (const std::exception&) {   // no exception crosses the C ABI
        end_streams();
        return gst_set_error(ctx, GST_INTERNAL, "learned batch: cannot start the selection thread");
    }
    bool engine_error = false;
    gst_status rs = gst_batch_run(ctx, next_job, job_done, (uint32_t)G->n, gmax, G->m, bopts, bstats, G->uid,
                                  std::max<int64_t>(valid, 1), &engine_error);
    {
        std::lock_guard<std::mutex> lk(mu);
        stop = true;   // an engine error ends the run early
    }
    cv.notify_all();
    sel_th.join();
    // selections still queued finish before their buffers are reused
    HIPCHK(hipStreamSynchronize((hipStream_t)ctx->stream2));
    if (lanes > 1) HIPCHK(hipStreamSynchronize((hipStream_t)ctx->stream3));
    end_streams();
    for (auto& x : qs) gst_result_free(&x.sub_res);
    // a query whose full solve never finished (the engine or the selection stopped early) carries
    // the error that stopped it, never a zeroed result marked OK
    const gst_status stop_err = engine_error ? rs : setup_err;
    if (stop_err != GST_OK)
        for (int32_t q = 0; q < nq; ++q)
            if (status[q] == GST_OK && !done[(size_t)q]) {
                gst_result_free(&results[q]);
                std::memset(&results[q], 0, sizeof(results[q]));
                results[q].root = -1;
                status[q] = stop_err;
            }
    if (setup_err != GST_OK) return setup_err;
    // a failed subgraph solve (no tree inside the subgraph) is no failure of its query: the
    // engine's own errors are returned, job outcomes are read from status[]
    if (engine_error) return rs;
    gst_status first = GST_OK;
    for (int32_t q = 0; q < nq && first == GST_OK; ++q)
        if (status[q] != GST_OK) first = status[q];
    return first;
}
