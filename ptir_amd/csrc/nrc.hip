// Neural radiance cache for gfx950: a multiresolution hash-grid encoding and a 64-wide MLP
// (four hidden ReLU layers, 3-channel exp output) whose products run on bf16 MFMA
// (v_mfma_f32_16x16x32_bf16) with FP32 accumulation.  Trained online on path suffixes that
// the tracer records at the first secondary vertex; queried there to end paths early.
//
// Operand layout.  All products are written transposed, samples as the N dimension:
//   forward   H_{l+1}^T = relu(W_l . H_l^T)    A = W_l      [out x in],   B = H_l^T  [in x samples]
//   backward  dH_l^T   = W_l^T . G_l^T         A = W_l^T    [in x out],   B = G_l^T  [out x samples]
//   weights   dW_l     = G_l^T . H_l           A = G_l^T    [out x smp],  B = H_l    [smp x in]
// A 16x16x32 MFMA gives lane (j = lane % 16, g = lane / 16) the outputs of sample j in rows
// 4g..4g+3 of its 16-row tile.  The K index of every product is permuted so that those
// rows are exactly what the lane has to supply as the B operand of the next product:
// slot kk of lane group g at k-step s holds neuron perm(s, g, kk) (two 16-row tiles give
// eight values).  Weights are packed once per update into per-lane fragments in that
// order; activations never leave registers between layers of the forward.
#include "internal.hpp"

namespace ptir {
namespace nrc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma(const bf16x8& a, const bf16x8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// neuron held by slot kk of lane group g at k-step s
__host__ __device__ constexpr int perm(int s, int g, int kk) {
    return 32 * s + (kk < 4 ? 4 * g + kk : 16 + 4 * g + (kk - 4));
}

__device__ __forceinline__ bf16x8 zero8() {
    bf16x8 z;
#pragma unroll
    for (int k = 0; k < 8; ++k) z[k] = (__bf16)0.f;
    return z;
}

// ---------------------------------------------------------------- encoding
struct Levels {
    int res[kLevels];
};

__device__ __forceinline__ uint32_t grid_index(int res, uint32_t x, uint32_t y, uint32_t z) {
    const uint32_t s = uint32_t(res) + 1u;
    if (uint64_t(s) * s * s <= uint64_t(kTableSize)) return x + y * s + z * s * s;
    return ((x * 1u) ^ (y * 2654435761u) ^ (z * 805459861u)) & uint32_t(kTableSize - 1);
}

// trilinear lookup of one level at p in [0,1]^3
__device__ __forceinline__ float2 grid_level(const float2* __restrict__ table, int level, int res, float px, float py,
                                             float pz) {
    const float ux = px * float(res), uy = py * float(res), uz = pz * float(res);
    const int ix = min(int(floorf(ux)), res - 1), iy = min(int(floorf(uy)), res - 1), iz = min(int(floorf(uz)), res - 1);
    const float fx = ux - float(ix), fy = uy - float(iy), fz = uz - float(iz);
    const float2* t = table + size_t(level) * kTableSize;
    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
        const float w = (cx ? fx : 1.f - fx) * (cy ? fy : 1.f - fy) * (cz ? fz : 1.f - fz);
        const float2 v = t[grid_index(res, uint32_t(ix + cx), uint32_t(iy + cy), uint32_t(iz + cz))];
        acc.x += w * v.x;
        acc.y += w * v.y;
    }
    return acc;
}

__device__ __forceinline__ void grid_level_bwd(float2* table_grad, int level, int res, float px, float py, float pz,
                                               float gx, float gy) {
    const float ux = px * float(res), uy = py * float(res), uz = pz * float(res);
    const int ix = min(int(floorf(ux)), res - 1), iy = min(int(floorf(uy)), res - 1), iz = min(int(floorf(uz)), res - 1);
    const float fx = ux - float(ix), fy = uy - float(iy), fz = uz - float(iz);
    float* t = reinterpret_cast<float*>(table_grad + size_t(level) * kTableSize);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
        const float w = (cx ? fx : 1.f - fx) * (cy ? fy : 1.f - fy) * (cz ? fz : 1.f - fz);
        const uint32_t i = grid_index(res, uint32_t(ix + cx), uint32_t(iy + cy), uint32_t(iz + cz));
        atomicAdd(t + 2 * i, w * gx);
        atomicAdd(t + 2 * i + 1, w * gy);
    }
}

__device__ __forceinline__ void unit_pos(const float* rec, const float* box, float& px, float& py, float& pz) {
    const float ex = fmaxf(box[3] - box[0], 1e-12f), ey = fmaxf(box[4] - box[1], 1e-12f),
                ez = fmaxf(box[5] - box[2], 1e-12f);
    px = fminf(fmaxf((rec[kRecPos] - box[0]) / ex, 0.f), 1.f);
    py = fminf(fmaxf((rec[kRecPos + 1] - box[1]) / ey, 0.f), 1.f);
    pz = fminf(fmaxf((rec[kRecPos + 2] - box[2]) / ez, 0.f), 1.f);
}

// real spherical harmonics of bands 0..3, coefficient q of unit direction d
__device__ __forceinline__ float sh4(int q, float x, float y, float z) {
    const float xx = x * x, yy = y * y, zz = z * z;
    switch (q) {
    case 0: return 0.28209479177387814f;
    case 1: return -0.48860251190291987f * y;
    case 2: return 0.48860251190291987f * z;
    case 3: return -0.48860251190291987f * x;
    case 4: return 1.0925484305920792f * x * y;
    case 5: return -1.0925484305920792f * y * z;
    case 6: return 0.94617469575755997f * zz - 0.31539156525251999f;
    case 7: return -1.0925484305920792f * x * z;
    case 8: return 0.54627421529603959f * (xx - yy);
    case 9: return 0.59004358992664352f * y * (3.f * xx - yy);
    case 10: return 2.8906114426405538f * x * y * z;
    case 11: return 0.45704579946446572f * y * (5.f * zz - 1.f);
    case 12: return 0.3731763325901154f * z * (5.f * zz - 3.f);
    case 13: return 0.45704579946446572f * x * (5.f * zz - 1.f);
    case 14: return 1.4453057213202769f * z * (xx - yy);
    default: return 0.59004358992664352f * x * (xx - 3.f * yy);
    }
}

// The eight input features lane group g supplies at k-step s for one query record:
//   s = 0: hash levels 2g, 2g+1 (slots 0..3) and 8+2g, 9+2g (slots 4..7)
//   s = 1: SH coefficients 4g..4g+3, then features 48+4g..: (roughness, albedo), (normal, 1), 0, 0
__device__ __forceinline__ bf16x8 encode_chunk(const float* rec, const float* box, const float2* table, const int* res,
                                               int g, int s) {
    float v[8];
    if (s == 0) {
        float px, py, pz;
        unit_pos(rec, box, px, py, pz);
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const int level = (h < 2 ? 2 * g : 8 + 2 * g) + (h & 1);
            const float2 f = grid_level(table, level, res[level], px, py, pz);
            v[2 * h] = f.x;
            v[2 * h + 1] = f.y;
        }
    } else {
        const float dx = rec[kRecDir], dy = rec[kRecDir + 1], dz = rec[kRecDir + 2];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = sh4(4 * g + k, dx, dy, dz);
        if (g == 0) {
            v[4] = rec[kRecRough];
            v[5] = rec[kRecAlb];
            v[6] = rec[kRecAlb + 1];
            v[7] = rec[kRecAlb + 2];
        } else if (g == 1) {
            v[4] = rec[kRecNrm] * 0.5f + 0.5f;
            v[5] = rec[kRecNrm + 1] * 0.5f + 0.5f;
            v[6] = rec[kRecNrm + 2] * 0.5f + 0.5f;
            v[7] = 1.f;
        } else {
            v[4] = v[5] = v[6] = v[7] = 0.f;
        }
    }
    bf16x8 o;
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = (__bf16)v[k];
    return o;
}

// ---------------------------------------------------------------- weight fragments
// Fragment f, lane l, slot kk (8 bf16 per lane, 16 B):
//   0..31   forward  W_l (l = f/8, row tile r, k-step s):  W_l[16r + i][perm(s,g,kk)]
//   32..33  forward  W_4 (k-step f-32):                    W_4[i][perm(s,g,kk)], rows >= 3 zero
//   34..65  backward W_l^T:                                W_l[perm(s,g,kk)][16r + i]
//   66..69  backward W_4^T (row tile f-66, K padded to 32): W_4[perm(0,g,kk)][16r + i], outputs >= 3 zero
constexpr int kFwdFrags = 34;
constexpr int kBwdFrags = 36;
static_assert(kFwdFrags + kBwdFrags == kFrags, "fragment count");

__global__ void k_pack(const float* __restrict__ w, bf16x8* __restrict__ frags) {
    const int t = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= kFrags * 64) return;
    const int f = t / 64, lane = t % 64, i = lane % 16, g = lane / 16;
    bf16x8 o;
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
        float v;
        if (f < 32) {
            const int l = f / 8, r = (f % 8) / 2, s = f % 2;
            v = w[l * 4096 + (16 * r + i) * 64 + perm(s, g, kk)];
        } else if (f < 34) {
            const int s = f - 32;
            v = i < kOut ? w[kOutOffset + i * 64 + perm(s, g, kk)] : 0.f;
        } else if (f < 66) {
            const int q = f - 34, l = q / 8, r = (q % 8) / 2, s = q % 2;
            v = w[l * 4096 + perm(s, g, kk) * 64 + 16 * r + i];
        } else {
            const int r = f - 66, o2 = perm(0, g, kk);
            v = o2 < kOut ? w[kOutOffset + o2 * 64 + 16 * r + i] : 0.f;
        }
        o[kk] = (__bf16)v;
    }
    frags[t] = o;
}

// ---------------------------------------------------------------- forward
struct FwdArgs {
    const float* rec;        // [n, kRecFloats]
    const float* box;        // [6] scene box lo.xyz hi.xyz
    const float2* table;     // [kLevels, kTableSize]
    const bf16x8* frags;     // [kFrags, 64]
    Levels lv;
    uint64_t n;
    // training: activations H_0..H_4 [5, n, 64] (slot order) and outputs z [n, 4]
    bf16x8* act;
    float4* z;
    // query: cache term thr * exp(z) added to pbr [n,3] and to the indirect part of comp [n,2,3];
    // pred [n,3] receives exp(z) (each optional)
    float* pbr;
    float* comp;
    float* pred;
};

constexpr int kBlock = 256;  // four waves, each a tile of 64 samples

template <bool TRAIN>
__global__ void __launch_bounds__(kBlock) k_forward(FwdArgs a) {
    __shared__ bf16x8 wl[kFwdFrags * 64];
    __shared__ int res[kLevels];
    for (int t = threadIdx.x; t < kFwdFrags * 64; t += kBlock) wl[t] = a.frags[t];
    if (threadIdx.x < kLevels) res[threadIdx.x] = a.lv.res[threadIdx.x];
    __syncthreads();
    const int lane = int(threadIdx.x) & 63, j = lane & 15, g = lane >> 4;
    const uint64_t tiles = (a.n + 63) / 64;
    for (uint64_t tile = uint64_t(blockIdx.x) * 4 + (threadIdx.x >> 6); tile < tiles; tile += uint64_t(gridDim.x) * 4) {
        bf16x8 B[4][2];
        bool live[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const uint64_t smp = tile * 64 + 16 * c + j;
            live[c] = smp < a.n;
            // records of paths that never reached the cache vertex (flag 0) carry no query:
            // they enter as zeros, whatever the rest of their record holds
            if (live[c] && a.rec[smp * kRecFloats + kRecFlag] != 0.f) {
                const float* rec = a.rec + smp * kRecFloats;
                B[c][0] = encode_chunk(rec, a.box, a.table, res, g, 0);
                B[c][1] = encode_chunk(rec, a.box, a.table, res, g, 1);
            } else {
                B[c][0] = B[c][1] = zero8();
            }
            if (TRAIN && live[c]) {
                bf16x8* h = a.act + smp * 8;
                h[g] = B[c][0];
                h[4 + g] = B[c][1];
            }
        }
#pragma unroll
        for (int l = 0; l < kHidden; ++l) {
            f32x4 acc[4][4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const bf16x8 w0 = wl[(l * 8 + r * 2) * 64 + lane], w1 = wl[(l * 8 + r * 2 + 1) * 64 + lane];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    acc[c][r] = mfma(w0, B[c][0], f32x4{0.f, 0.f, 0.f, 0.f});
                    acc[c][r] = mfma(w1, B[c][1], acc[c][r]);
                }
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
#pragma unroll
                for (int s = 0; s < 2; ++s)
#pragma unroll
                    for (int kk = 0; kk < 8; ++kk) B[c][s][kk] = (__bf16)fmaxf(acc[c][2 * s + kk / 4][kk % 4], 0.f);
                if (TRAIN && live[c]) {
                    bf16x8* h = a.act + (uint64_t(l + 1) * a.n + tile * 64 + 16 * c + j) * 8;
                    h[g] = B[c][0];
                    h[4 + g] = B[c][1];
                }
            }
        }
        const bf16x8 o0 = wl[32 * 64 + lane], o1 = wl[33 * 64 + lane];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            f32x4 z = mfma(o0, B[c][0], f32x4{0.f, 0.f, 0.f, 0.f});
            z = mfma(o1, B[c][1], z);
            const uint64_t smp = tile * 64 + 16 * c + j;
            if (g != 0 || !live[c]) continue;
            if (TRAIN) {
                a.z[smp] = make_float4(z[0], z[1], z[2], 0.f);
            } else {
                const float* rec = a.rec + smp * kRecFloats;
                const float p0 = expf(fminf(z[0], kZMax)), p1 = expf(fminf(z[1], kZMax)), p2 = expf(fminf(z[2], kZMax));
                if (rec[kRecFlag] != 0.f) {
                    if (a.pred) {
                        a.pred[3 * smp] = p0;
                        a.pred[3 * smp + 1] = p1;
                        a.pred[3 * smp + 2] = p2;
                    }
                    const float c0 = rec[kRecThr] * p0, c1 = rec[kRecThr + 1] * p1, c2 = rec[kRecThr + 2] * p2;
                    if (a.pbr) {
                        a.pbr[3 * smp] += c0;
                        a.pbr[3 * smp + 1] += c1;
                        a.pbr[3 * smp + 2] += c2;
                    }
                    if (a.comp) {
                        a.comp[6 * smp + 3] += c0;
                        a.comp[6 * smp + 4] += c1;
                        a.comp[6 * smp + 5] += c2;
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------- backward
struct BwdArgs {
    const float* rec;
    const bf16x8* frags;
    const bf16x8* act;       // [5, n, 8] fragments of 8 bf16 (slot order)
    const float4* z;         // [n]
    uint64_t n;
    float inv_n;             // loss normalisation
    float* dx;               // [n, 32] gradient of the hash features
    float* dw_part;          // [tiles, kParams] per-tile weight gradients
    float* loss;             // [1] sum of the per-sample losses / n (atomic)
};

constexpr int kStage = 72;  // bf16 per staging row: 64 samples + 8 pad (16 B aligned, spreads LDS banks)

// relative L2 gradient of one output channel
__device__ __forceinline__ float loss_grad(float z, float t, float inv_n, float& loss) {
    const float p = expf(fminf(z, kZMax));
    const float den = p * p + kLossEps;
    loss += (p - t) * (p - t) / den;
    return z > kZMax ? 0.f : 2.f * (p - t) * p / den * inv_n;
}

__global__ void __launch_bounds__(kBlock) k_backward(BwdArgs a) {
    __shared__ bf16x8 wl[kBwdFrags * 64];
    // per wave: G^T [64 out][64 samples], H^T [64 in][64 samples]; rows are read 16 B at a time
    __shared__ __attribute__((aligned(16))) __bf16 stage[4][2][64 * kStage];
    for (int t = threadIdx.x; t < kBwdFrags * 64; t += kBlock) wl[t] = a.frags[kFwdFrags * 64 + t];
    __syncthreads();
    const int wave = int(threadIdx.x) >> 6;
    const int lane = int(threadIdx.x) & 63, j = lane & 15, g = lane >> 4;
    __bf16* Gt = stage[wave][0];
    __bf16* Ht = stage[wave][1];
    const uint64_t tiles = (a.n + 63) / 64;
    const uint64_t tile = uint64_t(blockIdx.x) * 4 + wave;
    if (tile >= tiles) return;
    // loss gradient of the output layer; lanes of group 0 hold outputs 0..2 of their sample
    bf16x8 Bg[4][2];
    float lsum = 0.f;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const uint64_t smp = tile * 64 + 16 * c + j;
        Bg[c][0] = Bg[c][1] = zero8();
        if (g == 0 && smp < a.n) {
            const float* rec = a.rec + smp * kRecFloats;
            if (rec[kRecFlag] != 0.f) {
                const float4 z = a.z[smp];
                Bg[c][0][0] = (__bf16)loss_grad(z.x, rec[kRecTarget], a.inv_n, lsum);
                Bg[c][0][1] = (__bf16)loss_grad(z.y, rec[kRecTarget + 1], a.inv_n, lsum);
                Bg[c][0][2] = (__bf16)loss_grad(z.z, rec[kRecTarget + 2], a.inv_n, lsum);
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
    if (lane == 0 && a.loss) atomicAdd(a.loss, lsum * a.inv_n);
    float* part = a.dw_part + tile * kParams;
#pragma unroll
    for (int l = kHidden; l >= 0; --l) {
        const int rows = l == kHidden ? 16 : 64;  // output rows of this layer
        // stage G^T (from the B-layout registers) and H_l^T (from the stored activations)
        bf16x8 hv[4][2];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const uint64_t smp = tile * 64 + 16 * c + j;
            const int col = 16 * c + j;
            if (smp < a.n) {
                const bf16x8* h = a.act + (uint64_t(l) * a.n + smp) * 8;
                hv[c][0] = h[g];
                hv[c][1] = h[4 + g];
            } else {
                hv[c][0] = hv[c][1] = zero8();
            }
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int kk = 0; kk < 8; ++kk) {
                    const int nu = perm(s, g, kk);
                    Ht[nu * kStage + col] = hv[c][s][kk];
                    if (nu < rows) Gt[nu * kStage + col] = Bg[c][s][kk];
                }
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        // dW_l[out][in] = sum over the tile's samples of G[out] H[in]
        const int base = l == kHidden ? kOutOffset : l * 4096;
#pragma unroll
        for (int R = 0; R < rows / 16; ++R) {
            bf16x8 ga[2];
#pragma unroll
            for (int t = 0; t < 2; ++t)
                ga[t] = *reinterpret_cast<const bf16x8*>(Gt + (16 * R + j) * kStage + 32 * t + 8 * g);
#pragma unroll
            for (int N = 0; N < 4; ++N) {
                f32x4 d = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    const bf16x8 hb = *reinterpret_cast<const bf16x8*>(Ht + (16 * N + j) * kStage + 32 * t + 8 * g);
                    d = mfma(ga[t], hb, d);
                }
#pragma unroll
                for (int ii = 0; ii < 4; ++ii) part[base + (16 * R + 4 * g + ii) * 64 + 16 * N + j] = d[ii];
            }
        }
        __builtin_amdgcn_wave_barrier();
        // dH_l^T = W_l^T G^T: rows = in-neurons of layer l, K = out-neurons (32 for the output layer)
        const int fb = l == kHidden ? 32 : l * 8;  // first backward fragment of this layer
        const int ks = l == kHidden ? 1 : 2;
        f32x4 acc[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            bf16x8 w[2];
#pragma unroll
            for (int s = 0; s < 2; ++s) w[s] = s < ks ? wl[(fb + (l == kHidden ? r : r * 2 + s)) * 64 + lane] : zero8();
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                acc[c][r] = mfma(w[0], Bg[c][0], f32x4{0.f, 0.f, 0.f, 0.f});
                if (ks == 2) acc[c][r] = mfma(w[1], Bg[c][1], acc[c][r]);
            }
        }
        if (l == 0) {
            // gradient of the hash features (input rows 0..31)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const uint64_t smp = tile * 64 + 16 * c + j;
                if (smp >= a.n) continue;
                float* d = a.dx + smp * 32;
#pragma unroll
                for (int r = 0; r < 2; ++r)
#pragma unroll
                    for (int ii = 0; ii < 4; ++ii) d[16 * r + 4 * g + ii] = acc[c][r][ii];
            }
            break;
        }
        // through the ReLU of layer l-1: its output is H_l, held in hv in the same slots
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int kk = 0; kk < 8; ++kk) {
                    const float m = float(hv[c][s][kk]) > 0.f ? acc[c][2 * s + kk / 4][kk % 4] : 0.f;
                    Bg[c][s][kk] = (__bf16)m;
                }
    }
}

// dW = sum over tiles of the per-tile gradients, in two deterministic passes: blocks of
// kReduceChunk tiles first (blockIdx.y = chunk), then the chunk sums.
constexpr int kReduceChunk = 32;

__global__ void k_reduce_chunks(const float* __restrict__ part, uint64_t tiles, float* __restrict__ sums) {
    const int p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= kParams) return;
    const uint64_t t0 = uint64_t(blockIdx.y) * kReduceChunk;
    const uint64_t t1 = t0 + kReduceChunk < tiles ? t0 + kReduceChunk : tiles;
    float s = 0.f;
    for (uint64_t t = t0; t < t1; ++t) s += part[t * kParams + p];
    sums[uint64_t(blockIdx.y) * kParams + p] = s;
}

__global__ void k_reduce(const float* __restrict__ sums, uint64_t chunks, float* __restrict__ dw) {
    const int p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= kParams) return;
    float s = 0.f;
    for (uint64_t c = 0; c < chunks; ++c) s += sums[c * kParams + p];
    dw[p] = s;
}

__global__ void k_grid_bwd(const float* __restrict__ rec, const float* __restrict__ box, const float* __restrict__ dx,
                           Levels lv, uint64_t n, float2* table_grad) {
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n * kLevels) return;
    const uint64_t smp = t / kLevels;
    const int level = int(t % kLevels);
    const float* r = rec + smp * kRecFloats;
    if (r[kRecFlag] == 0.f) return;
    const float gx = dx[smp * 32 + 2 * level], gy = dx[smp * 32 + 2 * level + 1];
    if (gx == 0.f && gy == 0.f) return;
    float px, py, pz;
    unit_pos(r, box, px, py, pz);
    grid_level_bwd(table_grad, level, lv.res[level], px, py, pz, gx, gy);
}

// Adam with bias correction; the step counter lives on the device so a captured step replays.
// Each gradient is cleared after it is consumed.
__global__ void k_adam(float* __restrict__ p, float* __restrict__ grad, float* __restrict__ m, float* __restrict__ v,
                       uint64_t n, const uint32_t* __restrict__ step, AdamParams h) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float t = float(*step);
    const float c1 = 1.f / (1.f - powf(h.beta1, t)), c2 = 1.f / (1.f - powf(h.beta2, t));
    const float gr = grad[i];
    const float mi = h.beta1 * m[i] + (1.f - h.beta1) * gr;
    const float vi = h.beta2 * v[i] + (1.f - h.beta2) * gr * gr;
    m[i] = mi;
    v[i] = vi;
    p[i] -= h.lr * (mi * c1) / (sqrtf(vi * c2) + h.eps);
    grad[i] = 0.f;
}

__global__ void k_step_inc(uint32_t* step) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1u;
}

// counter-based uniform numbers for the initial parameters (PCG hash of index and seed)
__device__ __forceinline__ float hash_uniform(uint64_t i, uint32_t seed) {
    uint32_t s = uint32_t(i) * 747796405u + 2891336453u + seed * 2654435761u + uint32_t(i >> 32) * 40503u;
    s = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    s = (s >> 22u) ^ s;
    return float(s >> 8) * (1.0f / 16777216.0f);
}

__global__ void k_init(float* __restrict__ w, float* __restrict__ table, uint64_t table_n, uint32_t seed) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < uint64_t(kParams)) {
        float v = 0.f;
        const int p = int(i);
        if (p < kOutOffset) v = (2.f * hash_uniform(i, seed) - 1.f) * sqrtf(6.f / 64.f);
        else if ((p - kOutOffset) / 64 < kOut) v = (2.f * hash_uniform(i, seed) - 1.f) * sqrtf(6.f / 67.f);
        w[p] = v;
    }
    if (i < table_n) table[i] = (2.f * hash_uniform(i, seed ^ 0x9e3779b9u) - 1.f) * 1e-4f;
}

}  // namespace nrc

// ---------------------------------------------------------------- host side
namespace {
template <class T>
hipError_t dalloc(T*& p, size_t count) {
    return hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
}
}  // namespace

hipError_t nrc_alloc(NrcBuffers& b, uint64_t max_batch, uint32_t seed, hipStream_t s) {
    using namespace nrc;
    b.max_batch = max_batch;
    const uint64_t tn = nrc_table_floats();
    const uint64_t tiles = (max_batch + 63) / 64;
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = dalloc(b.w, kParams);
    if (e == hipSuccess) e = dalloc(b.dw, kParams);
    if (e == hipSuccess) e = dalloc(b.mw, kParams);
    if (e == hipSuccess) e = dalloc(b.vw, kParams);
    if (e == hipSuccess) e = dalloc(b.table, tn);
    if (e == hipSuccess) e = dalloc(b.dtable, tn);
    if (e == hipSuccess) e = dalloc(b.mt, tn);
    if (e == hipSuccess) e = dalloc(b.vt, tn);
    if (e == hipSuccess) e = hipMalloc(&b.frags, size_t(kFrags) * 64 * 16);
    if (e == hipSuccess) e = hipMalloc(&b.act, size_t(kHidden + 1) * max_batch * 128);
    if (e == hipSuccess) e = dalloc(b.z, max_batch * 4);
    if (e == hipSuccess) e = dalloc(b.dx, max_batch * 32);
    if (e == hipSuccess) e = dalloc(b.dw_part, tiles * kParams);
    if (e == hipSuccess) e = dalloc(b.dw_chunks, ((tiles + kReduceChunk - 1) / kReduceChunk) * kParams);
    if (e == hipSuccess) e = dalloc(b.loss, 1);
    if (e == hipSuccess) e = dalloc(b.step, 1);
    if (e != hipSuccess) return e;
    e = fill_zero(b.dw, kParams * sizeof(float), s);
    if (e == hipSuccess) e = fill_zero(b.mw, kParams * sizeof(float), s);
    if (e == hipSuccess) e = fill_zero(b.vw, kParams * sizeof(float), s);
    if (e == hipSuccess) e = fill_zero(b.dtable, tn * sizeof(float), s);
    if (e == hipSuccess) e = fill_zero(b.mt, tn * sizeof(float), s);
    if (e == hipSuccess) e = fill_zero(b.vt, tn * sizeof(float), s);
    if (e == hipSuccess) e = fill_zero(b.loss, sizeof(float), s);
    if (e == hipSuccess) e = fill_zero(b.step, sizeof(uint32_t), s);
    if (e != hipSuccess) return e;
    k_init<<<unsigned((tn + 255) / 256), 256, 0, s>>>(b.w, b.table, tn, seed);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return nrc_pack(b, s);
}

void nrc_free(NrcBuffers& b) {
    void* ps[] = {b.w, b.dw, b.mw, b.vw, b.table, b.dtable, b.mt, b.vt, b.frags, b.act, b.z, b.dx, b.dw_part,
                  b.dw_chunks, b.loss, b.step};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    b = NrcBuffers{};
}

hipError_t nrc_pack(NrcBuffers& b, hipStream_t s) {
    nrc::k_pack<<<(nrc::kFrags * 64 + 255) / 256, 256, 0, s>>>(b.w, static_cast<nrc::bf16x8*>(b.frags));
    return hipGetLastError();
}

static nrc::Levels levels() {
    nrc::Levels lv;
    for (int l = 0; l < nrc::kLevels; ++l) lv.res[l] = nrc_level_res(l);
    return lv;
}

static unsigned fwd_grid(uint64_t n) {
    const uint64_t blocks = (n + 255) / 256;
    return unsigned(blocks < 2048 ? blocks : 2048);
}

hipError_t nrc_query(const NrcBuffers& b, const float* rec, uint64_t n, const float* box, float* pbr, float* comp,
                     float* pred, hipStream_t s) {
    if (n == 0) return hipSuccess;
    nrc::FwdArgs a{rec, box, reinterpret_cast<const float2*>(b.table), static_cast<const nrc::bf16x8*>(b.frags),
                   levels(), n, nullptr, nullptr, pbr, comp, pred};
    nrc::k_forward<false><<<fwd_grid(n), nrc::kBlock, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t nrc_backprop(NrcBuffers& b, const float* rec, uint64_t n, const float* box, hipStream_t s) {
    using namespace nrc;
    if (n == 0) return hipSuccess;
    hipError_t e = fill_zero(b.loss, sizeof(float), s);
    if (e != hipSuccess) return e;
    FwdArgs f{rec, box, reinterpret_cast<const float2*>(b.table), static_cast<const bf16x8*>(b.frags), levels(), n,
              static_cast<bf16x8*>(b.act), reinterpret_cast<float4*>(b.z), nullptr, nullptr, nullptr};
    k_forward<true><<<fwd_grid(n), kBlock, 0, s>>>(f);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const uint64_t tiles = (n + 63) / 64;
    BwdArgs g{rec, static_cast<const bf16x8*>(b.frags), static_cast<const bf16x8*>(b.act),
              reinterpret_cast<const float4*>(b.z), n, 1.f / float(n), b.dx, b.dw_part, b.loss};
    k_backward<<<unsigned((tiles + 3) / 4), kBlock, 0, s>>>(g);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const uint64_t chunks = (tiles + kReduceChunk - 1) / kReduceChunk;
    k_reduce_chunks<<<dim3((kParams + 255) / 256, unsigned(chunks)), 256, 0, s>>>(b.dw_part, tiles, b.dw_chunks);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_reduce<<<(kParams + 255) / 256, 256, 0, s>>>(b.dw_chunks, chunks, b.dw);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_grid_bwd<<<unsigned((n * kLevels + 255) / 256), 256, 0, s>>>(rec, box, b.dx, levels(), n,
                                                                    reinterpret_cast<float2*>(b.dtable));
    return hipGetLastError();
}

hipError_t nrc_step(NrcBuffers& b, const nrc::AdamParams& h, hipStream_t s) {
    using namespace nrc;
    k_step_inc<<<1, 64, 0, s>>>(b.step);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_adam<<<(kParams + 255) / 256, 256, 0, s>>>(b.w, b.dw, b.mw, b.vw, kParams, b.step, h);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const uint64_t tn = nrc_table_floats();
    k_adam<<<unsigned((tn + 255) / 256), 256, 0, s>>>(b.table, b.dtable, b.mt, b.vt, tn, b.step, h);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return nrc_pack(b, s);
}

}  // namespace ptir
