// Learned stand-in for the sub-stepped 4TM + LLB film integrator, gfx950.
//
// The network is an MLP 11 -> 64 -> 64 -> 6 with SiLU.  Operands are BF16 on
// v_mfma_f32_16x16x16_bf16 with FP32 accumulation.  The product is written transposed,
// Z^T = W * X^T: weights are the A operand and activations the B operand, cells or
// samples in the 16 columns.  For this instruction a lane l holds
//   A[l%16][4(l/16)+v], B[4(l/16)+v][l%16], D[4(l/16)+v][l%16]   (v = 0..3),
// so a layer's D registers are already the next layer's B operand and inference needs
// no data exchange between layers.  Weight gradients reduce over samples (the K
// dimension); their operands go through LDS in [feature][sample] order, where both
// A and B fragments are contiguous 8-byte reads.
//
// It stands in for the substep loop of multiphysics_kernel! (src/fdtd/Kernels.jl:362-412);
// the plain substep it is checked against, and falls back to, is the same llb_step +
// update_4tm code the plain integrator runs (film_physics.h).
#include <algorithm>
#include <cmath>

#include "film_physics.h"

namespace mp {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int OW1 = 0, OB1 = 1024, OW2 = 1088, OB2 = 5184, OW3 = 5248, OB3 = 6272;
constexpr int NPARAM = MP_MLP_NPARAM;
// blocks of the surrogate film update (4 waves each): one wave per SIMD on 256 CUs
#ifndef SUR_BLOCKS
#define SUR_BLOCKS 256
#endif
// fragment index ranges inside the packed buffer (64 lanes x uint2 each)
constexpr int F_W1 = 0, F_W2 = 4, F_W3 = 20, F_W3T = 24, F_W2T = 28, NFRAG = 44, NFWD = 24;
static_assert(NFRAG * 64 * 8 == MP_MLP_PACKED_BYTES, "packed size");
// norm buffer offsets
constexpr int N_XM = 0, N_XIS = 16, N_YM = 32, N_YS = 48, N_YIS = 64;

__device__ __forceinline__ v4f mfma16(v4s a, v4s b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ uint16_t bfbits(float x) { return st<bf16>(x).v; }
__device__ __forceinline__ v4s bf4(v4f x) {
    v4s r;
#pragma unroll
    for (int v = 0; v < 4; ++v) r[v] = (short)bfbits(x[v]);
    return r;
}
__device__ __forceinline__ v4s as_v4s(uint2 u) {
    v4s r;
    r[0] = (short)(u.x & 0xffffu);
    r[1] = (short)(u.x >> 16);
    r[2] = (short)(u.y & 0xffffu);
    r[3] = (short)(u.y >> 16);
    return r;
}
// Logistic function for the SiLU activations, on the native exp2 and reciprocal
// instructions (about 1 ulp each): the result is rounded to BF16 before it is used as
// an operand, so a correctly rounded exp and division would buy nothing.
__device__ __forceinline__ float sigm(float z) {
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.44269504088896341f * z));
}

// Raw features of one cell (the order of mp_mlp in mphip.h).
__device__ __forceinline__ void cell_features(float f[MP_MLP_IN], double te, double tl, double tsT,
                                              double tsR, double mT, double mR, double pabs, double dt,
                                              const FilmConsts &gd) {
    const double pn = pabs * 1e-18;
    const double apn = __builtin_fabs(pn);
    const double ash = __builtin_copysign(log(apn + sqrt(apn * apn + 1.0)), pn);
    f[0] = (float)te;
    f[1] = (float)tl;
    f[2] = (float)tsT;
    f[3] = (float)tsR;
    f[4] = (float)mT;
    f[5] = (float)mR;
    f[6] = (float)ash;
    f[7] = (float)log(dt * 1e18);
    f[8] = (float)gd.gamma_e;
    f[9] = (float)log(gd.Gel);
    f[10] = (float)gd.T_Curie;
}

// One plain substep of the film integrator (an iteration of the multiphysics_kernel loop).
template <typename L>
__device__ __forceinline__ void plain_substep(double &te, double &tl, double &tsT, double &tsR, double m[6],
                                              double pabs, const FilmConsts &gd, const L *tm_lut,
                                              const L *re_lut, double T_min, double inv_dT, int32_t lut_N,
                                              double sub_dt, int iters) {
    double out[4];
    llb_step<L>(m, te, tsT, tsR, gd, tm_lut, re_lut, T_min, inv_dT, lut_N, sub_dt, iters, out);
    const double Q_spin_TM = gd.Ms_TM * out[2] * (out[0] / sub_dt);
    const double Q_spin_RE = gd.Ms_RE * out[3] * (out[1] / sub_dt);
    update_4tm(te, tl, tsT, tsR, pabs, Q_spin_TM, Q_spin_RE, gd, sub_dt);
}

// ---- packing -------------------------------------------------------------------------

__global__ __launch_bounds__(64) void mlp_pack_kernel(const float *__restrict__ P, uint2 *__restrict__ out) {
    const int f = blockIdx.x, l = threadIdx.x;
    const int r = l % 16, q = 4 * (l / 16);
    float w[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        if (f < F_W2) {
            w[v] = P[OW1 + (16 * f + r) * 16 + q + v];
        } else if (f < F_W3) {
            const int mt = (f - F_W2) / 4, kt = (f - F_W2) % 4;
            w[v] = P[OW2 + (16 * mt + r) * 64 + 16 * kt + q + v];
        } else if (f < F_W3T) {
            const int kt = f - F_W3;
            w[v] = P[OW3 + r * 64 + 16 * kt + q + v];
        } else if (f < F_W2T) {
            const int mt = f - F_W3T;  // W3^T[16mt + r][q + v] = W3[q + v][16mt + r]
            w[v] = P[OW3 + (q + v) * 64 + 16 * mt + r];
        } else {
            const int mt = (f - F_W2T) / 4, kt = (f - F_W2T) % 4;  // W2^T
            w[v] = P[OW2 + (16 * kt + q + v) * 64 + 16 * mt + r];
        }
    }
    out[f * 64 + l] = make_uint2((uint32_t)bfbits(w[0]) | ((uint32_t)bfbits(w[1]) << 16),
                                 (uint32_t)bfbits(w[2]) | ((uint32_t)bfbits(w[3]) << 16));
}

// ---- inference -----------------------------------------------------------------------

// Shared block state for inference: the 24 forward fragments and the biases.
struct InferLds {
    uint2 W[NFWD * 64];
    float b[144];  // b1[64], b2[64], b3[16]
};

__device__ __forceinline__ void load_infer_lds(InferLds &s, const uint2 *__restrict__ packed,
                                               const float *__restrict__ P) {
    for (int i = threadIdx.x; i < NFWD * 64; i += blockDim.x) s.W[i] = packed[i];
    for (int i = threadIdx.x; i < 144; i += blockDim.x)
        s.b[i] = i < 64 ? P[OB1 + i] : (i < 128 ? P[OB2 + i - 64] : P[OB3 + i - 128]);
}

// Forward of one 16-column tile: xb is the B operand of the (normalised, BF16) features;
// returns the output fragment (rows = outputs 4(l/16)+v, column = l%16).
__device__ __forceinline__ v4f mlp_tile_forward(const InferLds &s, v4s xb, int l) {
    const int q = 4 * (l / 16);
    v4s h1[4], h2[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        v4f c = {s.b[16 * mt + q], s.b[16 * mt + q + 1], s.b[16 * mt + q + 2], s.b[16 * mt + q + 3]};
        c = mfma16(as_v4s(s.W[(F_W1 + mt) * 64 + l]), xb, c);
        v4f h;
#pragma unroll
        for (int v = 0; v < 4; ++v) h[v] = c[v] * sigm(c[v]);
        h1[mt] = bf4(h);
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        v4f c = {s.b[64 + 16 * mt + q], s.b[64 + 16 * mt + q + 1], s.b[64 + 16 * mt + q + 2],
                 s.b[64 + 16 * mt + q + 3]};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) c = mfma16(as_v4s(s.W[(F_W2 + mt * 4 + kt) * 64 + l]), h1[kt], c);
        v4f h;
#pragma unroll
        for (int v = 0; v < 4; ++v) h[v] = c[v] * sigm(c[v]);
        h2[mt] = bf4(h);
    }
    v4f y = {s.b[128 + q], s.b[128 + q + 1], s.b[128 + q + 2], s.b[128 + q + 3]};
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) y = mfma16(as_v4s(s.W[(F_W3 + kt) * 64 + l]), h2[kt], y);
    return y;
}

// Normalised BF16 features of one row into xs[16] (row padded with zeros).
__device__ __forceinline__ void stage_features(uint16_t *xs, const float f[MP_MLP_IN], const float *__restrict__ norm) {
#pragma unroll
    for (int i = 0; i < 16; ++i)
        xs[i] = i < MP_MLP_IN ? bfbits((f[i] - norm[N_XM + i]) * norm[N_XIS + i]) : (uint16_t)0;
}

// 64 rows per wave: each lane stages its row's features, the wave runs 4 tiles, and
// each lane reads back its row's 6 raw rates.
__device__ __forceinline__ void wave_infer(const InferLds &s, uint16_t (*xs)[16], float (*ys)[8],
                                           const float *__restrict__ norm, float rates[MP_MLP_OUT]) {
    const int l = threadIdx.x & 63;
    const int q = 4 * (l / 16);
    __syncthreads();
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const uint2 u = *reinterpret_cast<const uint2 *>(&xs[16 * nt + l % 16][q]);
        const v4f y = mlp_tile_forward(s, as_v4s(u), l);
        if (q < 8) {
#pragma unroll
            for (int v = 0; v < 4; ++v) ys[16 * nt + l % 16][q + v] = y[v];
        }
    }
    __syncthreads();
#pragma unroll
    for (int o = 0; o < MP_MLP_OUT; ++o) rates[o] = ys[l][o] * norm[N_YS + o] + norm[N_YM + o];
}

__global__ __launch_bounds__(256) void mlp_forward_kernel(const uint2 *__restrict__ packed,
                                                          const float *__restrict__ P,
                                                          const float *__restrict__ norm,
                                                          const float *__restrict__ X, float *__restrict__ Y,
                                                          int64_t B) {
    __shared__ InferLds s;
    __shared__ __align__(16) uint16_t xs[4][64][16];
    __shared__ float ys[4][64][8];
    load_infer_lds(s, packed, P);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
    float f[MP_MLP_IN];
#pragma unroll
    for (int i = 0; i < MP_MLP_IN; ++i) f[i] = b < B ? X[b * MP_MLP_IN + i] : norm[N_XM + i];
    stage_features(xs[w][l], f, norm);
    float r[MP_MLP_OUT];
    wave_infer(s, xs[w], ys[w], norm, r);
    if (b < B) {
#pragma unroll
        for (int o = 0; o < MP_MLP_OUT; ++o) Y[b * MP_MLP_OUT + o] = r[o];
    }
}

// ---- training: forward + backward ------------------------------------------------------

struct TrainLds {
    uint2 W[NFRAG * 64];
    float b[144];
};
struct WaveLds {  // [feature][sample] BF16 tiles of one 16-sample tile
    uint16_t x[16][16], dy[16][16];
    uint16_t h1[64][16], h2[64][16], dz1[64][16], dz2[64][16];
};

constexpr int TRAIN_BLOCKS_MAX = 256;
constexpr int PART_STRIDE = NPARAM + 1;  // + loss

__global__ __launch_bounds__(256) void mlp_train_kernel(const uint2 *__restrict__ packed, const float *__restrict__ P,
                                                        const float *__restrict__ norm, const float *__restrict__ X,
                                                        const float *__restrict__ T, int64_t B, int64_t tiles_per_wave,
                                                        float gscale, float *__restrict__ part) {
    __shared__ TrainLds s;
    __shared__ __align__(16) WaveLds wl[4];
    for (int i = threadIdx.x; i < NFRAG * 64; i += blockDim.x) s.W[i] = packed[i];
    for (int i = threadIdx.x; i < 144; i += blockDim.x)
        s.b[i] = i < 64 ? P[OB1 + i] : (i < 128 ? P[OB2 + i - 64] : P[OB3 + i - 128]);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int c = l % 16, q = 4 * (l / 16);
    const int64_t wave_id = (int64_t)blockIdx.x * 4 + w;
    const int64_t nwaves = (int64_t)gridDim.x * 4;
    WaveLds &L = wl[w];

    v4f gW1[4], gW2[4][4], gW3[4];
    float gb1[4][4], gb2[4][4], gb3[4];
    float loss = 0.f;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        gW1[a] = v4f{0.f, 0.f, 0.f, 0.f};
        gW3[a] = v4f{0.f, 0.f, 0.f, 0.f};
        gb3[a] = 0.f;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            gW2[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
            gb1[a][b] = 0.f;
            gb2[a][b] = 0.f;
        }
    }
    __syncthreads();

    for (int64_t it = 0; it < tiles_per_wave; ++it) {
        const int64_t tile = wave_id + it * nwaves;
        const int64_t row = tile * 16 + c;
        const bool valid = row < B;
        // features q..q+3 of sample c (B operand) and the [feature][sample] LDS copy
        v4f xf;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int fi = q + v;
            xf[v] = (valid && fi < MP_MLP_IN) ? (X[row * MP_MLP_IN + fi] - norm[N_XM + fi]) * norm[N_XIS + fi] : 0.f;
        }
        const v4s xb = bf4(xf);
#pragma unroll
        for (int v = 0; v < 4; ++v) L.x[q + v][c] = (uint16_t)xb[v];

        // forward, activations kept in registers
        v4f z1[4], z2[4];
        v4s h1[4], h2[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            v4f a = {s.b[16 * mt + q], s.b[16 * mt + q + 1], s.b[16 * mt + q + 2], s.b[16 * mt + q + 3]};
            z1[mt] = mfma16(as_v4s(s.W[(F_W1 + mt) * 64 + l]), xb, a);
            v4f h;
#pragma unroll
            for (int v = 0; v < 4; ++v) h[v] = z1[mt][v] * sigm(z1[mt][v]);
            h1[mt] = bf4(h);
#pragma unroll
            for (int v = 0; v < 4; ++v) L.h1[16 * mt + q + v][c] = (uint16_t)h1[mt][v];
        }
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            v4f a = {s.b[64 + 16 * mt + q], s.b[64 + 16 * mt + q + 1], s.b[64 + 16 * mt + q + 2],
                     s.b[64 + 16 * mt + q + 3]};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) a = mfma16(as_v4s(s.W[(F_W2 + mt * 4 + kt) * 64 + l]), h1[kt], a);
            z2[mt] = a;
            v4f h;
#pragma unroll
            for (int v = 0; v < 4; ++v) h[v] = a[v] * sigm(a[v]);
            h2[mt] = bf4(h);
#pragma unroll
            for (int v = 0; v < 4; ++v) L.h2[16 * mt + q + v][c] = (uint16_t)h2[mt][v];
        }
        v4f y = {s.b[128 + q], s.b[128 + q + 1], s.b[128 + q + 2], s.b[128 + q + 3]};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) y = mfma16(as_v4s(s.W[(F_W3 + kt) * 64 + l]), h2[kt], y);

        // loss and dL/dy
        v4f dy;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int o = q + v;
            if (valid && o < MP_MLP_OUT) {
                const float tn = (T[row * MP_MLP_OUT + o] - norm[N_YM + o]) * norm[N_YIS + o];
                const float e = y[v] - tn;
                loss += e * e;
                dy[v] = e * gscale;
            } else {
                dy[v] = 0.f;
            }
            gb3[v] += dy[v];
        }
        const v4s dyb = bf4(dy);
#pragma unroll
        for (int v = 0; v < 4; ++v) L.dy[q + v][c] = (uint16_t)dyb[v];

        // backward through the layers (weights transposed as A operands)
        v4s dz2b[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            v4f d = mfma16(as_v4s(s.W[(F_W3T + mt) * 64 + l]), dyb, v4f{0.f, 0.f, 0.f, 0.f});
            v4f dz;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float sg = sigm(z2[mt][v]);
                dz[v] = d[v] * (sg * (1.0f + z2[mt][v] * (1.0f - sg)));
                gb2[mt][v] += dz[v];
            }
            dz2b[mt] = bf4(dz);
#pragma unroll
            for (int v = 0; v < 4; ++v) L.dz2[16 * mt + q + v][c] = (uint16_t)dz2b[mt][v];
        }
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            v4f d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) d = mfma16(as_v4s(s.W[(F_W2T + mt * 4 + kt) * 64 + l]), dz2b[kt], d);
            v4f dz;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float sg = sigm(z1[mt][v]);
                dz[v] = d[v] * (sg * (1.0f + z1[mt][v] * (1.0f - sg)));
                gb1[mt][v] += dz[v];
            }
            const v4s dzb = bf4(dz);
#pragma unroll
            for (int v = 0; v < 4; ++v) L.dz1[16 * mt + q + v][c] = (uint16_t)dzb[v];
        }
        __syncthreads();

        // weight gradients: K = the 16 samples of this tile
        const v4s a_dy = as_v4s(*reinterpret_cast<const uint2 *>(&L.dy[c][q]));
        const v4s b_x = as_v4s(*reinterpret_cast<const uint2 *>(&L.x[c][q]));
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const v4s bh2 = as_v4s(*reinterpret_cast<const uint2 *>(&L.h2[16 * nt + c][q]));
            gW3[nt] = mfma16(a_dy, bh2, gW3[nt]);
        }
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const v4s adz2 = as_v4s(*reinterpret_cast<const uint2 *>(&L.dz2[16 * mt + c][q]));
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const v4s bh1 = as_v4s(*reinterpret_cast<const uint2 *>(&L.h1[16 * nt + c][q]));
                gW2[mt][nt] = mfma16(adz2, bh1, gW2[mt][nt]);
            }
            const v4s adz1 = as_v4s(*reinterpret_cast<const uint2 *>(&L.dz1[16 * mt + c][q]));
            gW1[mt] = mfma16(adz1, b_x, gW1[mt]);
        }
        __syncthreads();
    }

    // bias sums over the 16 sample columns (lanes sharing l/16), fixed order
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
#pragma unroll
        for (int a = 0; a < 4; ++a) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                gb1[a][v] += __shfl_xor(gb1[a][v], off);
                gb2[a][v] += __shfl_xor(gb2[a][v], off);
            }
            gb3[a] += __shfl_xor(gb3[a], off);
        }
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) loss += __shfl_xor(loss, off);

    float *dst = part + wave_id * PART_STRIDE;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            dst[OW1 + (16 * mt + q + v) * 16 + c] = gW1[mt][v];
            dst[OW3 + (q + v) * 64 + 16 * mt + c] = gW3[mt][v];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) dst[OW2 + (16 * mt + q + v) * 64 + 16 * nt + c] = gW2[mt][nt][v];
        }
    }
    if (c == 0) {
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                dst[OB1 + 16 * mt + q + v] = gb1[mt][v];
                dst[OB2 + 16 * mt + q + v] = gb2[mt][v];
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) dst[OB3 + q + v] = gb3[v];
    }
    if (l == 0) dst[NPARAM] = loss;
}

__global__ __launch_bounds__(256) void mlp_reduce_kernel(const float *__restrict__ part, int64_t nwaves,
                                                         float *__restrict__ grad, double *__restrict__ loss,
                                                         double loss_scale) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > NPARAM) return;
    double s = 0.0;
    for (int64_t w = 0; w < nwaves; ++w) s += (double)part[w * PART_STRIDE + i];
    if (i < NPARAM) grad[i] = (float)s;
    else if (loss) *loss = s * loss_scale;
}

__global__ __launch_bounds__(256) void mlp_adam_kernel(float *__restrict__ P, const float *__restrict__ g,
                                                       float *__restrict__ m1, float *__restrict__ m2, float lr,
                                                       float b1, float b2, float eps, float wd, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float b = b2 * m2[i] + (1.0f - b2) * (gi * gi);
    m1[i] = a;
    m2[i] = b;
    const float mh = a / c1, vh = b / c2;
    P[i] = P[i] - lr * (mh / (sqrtf(vh) + eps) + wd * P[i]);
}

// ---- dataset and the surrogate step -----------------------------------------------------

template <typename L, bool HAS_PABS>
__global__ __launch_bounds__(128) void dataset_kernel(
    const L *__restrict__ Te, const L *__restrict__ Tl, const L *__restrict__ TsT, const L *__restrict__ TsR,
    const L *__restrict__ mTx, const L *__restrict__ mTy, const L *__restrict__ mTz, const L *__restrict__ mRx,
    const L *__restrict__ mRy, const L *__restrict__ mRz, const double *__restrict__ window,
    const FilmConsts *__restrict__ gdp, const L *__restrict__ tm_lut, const L *__restrict__ re_lut, double T_min,
    double inv_dT, int32_t lut_N, double dt, int subcycles, int iters, int64_t stride, int64_t rows,
    float *__restrict__ X, float *__restrict__ T) {
    const FilmConsts &gd = *gdp;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int64_t n = r * stride;
    const double pabs = HAS_PABS ? window[n] / dt : 0.0;
    double te = (double)Te[n], tl = (double)Tl[n], tsT = (double)TsT[n], tsR = (double)TsR[n];
    double m[6] = {(double)mTx[n], (double)mTy[n], (double)mTz[n], (double)mRx[n], (double)mRy[n], (double)mRz[n]};
    float f[MP_MLP_IN];
    cell_features(f, te, tl, tsT, tsR, m[0], m[3], pabs, dt, gd);
#pragma unroll
    for (int i = 0; i < MP_MLP_IN; ++i) X[r * MP_MLP_IN + i] = f[i];
    const double te0 = te, tl0 = tl, tsT0 = tsT, tsR0 = tsR, mT0 = m[0], mR0 = m[3];
    const double sub_dt = dt / (double)subcycles;
    for (int s = 0; s < subcycles; ++s)
        plain_substep<L>(te, tl, tsT, tsR, m, pabs, gd, tm_lut, re_lut, T_min, inv_dT, lut_N, sub_dt, iters);
    const double hg = 0.5 * gd.gamma_e;
    float *t = T + r * MP_MLP_OUT;
    t[0] = (float)((hg * (te * te) - hg * (te0 * te0)) / dt - pabs);
    t[1] = (float)((tl - tl0) / dt);
    t[2] = (float)((tsT - tsT0) / dt);
    t[3] = (float)((tsR - tsR0) / dt);
    t[4] = (float)((m[0] - mT0) / dt);
    t[5] = (float)((m[3] - mR0) / dt);
}

// A cell that failed the check: its start state (FP64, as the plain loop holds it), its
// absorbed power and its index.  Fallback cells are compacted into a list and take every
// plain substep in a second kernel, so the substep loop runs on full waves instead of
// diverging inside the first.
struct FallbackRec {
    double s[11];  // te, tl, tsT, tsR, m[6], pabs
    int64_t cell;
};

__global__ void reset_count_kernel(unsigned long long *c) { *c = 0ull; }

template <typename L>
__global__ __launch_bounds__(128) void surrogate_fallback_kernel(
    L *__restrict__ Te, L *__restrict__ Tl, L *__restrict__ TsT, L *__restrict__ TsR, L *__restrict__ mTx,
    L *__restrict__ mTy, L *__restrict__ mTz, L *__restrict__ mRx, L *__restrict__ mRy, L *__restrict__ mRz,
    const FallbackRec *__restrict__ recs, const unsigned long long *__restrict__ nrec, const FilmConsts *__restrict__ gdp,
    const L *__restrict__ tm_lut, const L *__restrict__ re_lut, double T_min, double inv_dT, int32_t lut_N,
    double sub_dt, int subcycles, int iters) {
    const FilmConsts &gd = *gdp;
    const unsigned long long cnt = *nrec;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < cnt;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        const FallbackRec &r = recs[i];
        double te = r.s[0], tl = r.s[1], tsT = r.s[2], tsR = r.s[3];
        double m[6] = {r.s[4], r.s[5], r.s[6], r.s[7], r.s[8], r.s[9]};
        const double pabs = r.s[10];
        const int64_t n = r.cell;
        for (int k = 0; k < subcycles; ++k)
            plain_substep<L>(te, tl, tsT, tsR, m, pabs, gd, tm_lut, re_lut, T_min, inv_dT, lut_N, sub_dt, iters);
        Te[n] = (L)te; Tl[n] = (L)tl; TsT[n] = (L)tsT; TsR[n] = (L)tsR;
        mTx[n] = (L)m[0]; mTy[n] = (L)m[1]; mTz[n] = (L)m[2];
        mRx[n] = (L)m[3]; mRy[n] = (L)m[4]; mRz[n] = (L)m[5];
    }
}

template <typename L, bool HAS_PABS>
__global__ __launch_bounds__(256) void surrogate_kernel(
    L *__restrict__ Te, L *__restrict__ Tl, L *__restrict__ TsT, L *__restrict__ TsR, L *__restrict__ mTx,
    L *__restrict__ mTy, L *__restrict__ mTz, L *__restrict__ mRx, L *__restrict__ mRy, L *__restrict__ mRz,
    double *__restrict__ window, const FilmConsts *__restrict__ gdp, const L *__restrict__ tm_lut,
    const L *__restrict__ re_lut, double T_min, double inv_dT, int32_t lut_N, double dt, int subcycles,
    int iters, int64_t N, const uint2 *__restrict__ packed, const float *__restrict__ P,
    const float *__restrict__ norm, double tol, unsigned long long *__restrict__ fallbacks,
    FallbackRec *__restrict__ recs, unsigned long long *__restrict__ nrec) {
    const FilmConsts &gd = *gdp;
    __shared__ InferLds s;
    __shared__ __align__(16) uint16_t xs[4][64][16];
    __shared__ float ys[4][64][8];
    load_infer_lds(s, packed, P);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    // the blocks loop over the cells 256 at a time (a lane past the end only occurs in a
    // block's last pass, so its early return skips no later barrier)
    for (int64_t base = (int64_t)blockIdx.x * 256; base < N; base += (int64_t)gridDim.x * 256) {
    const int64_t n = base + threadIdx.x;
    const bool active = n < N;

    double te = 300.0, tl = 300.0, tsT = 300.0, tsR = 300.0, pabs = 0.0;
    double m[6] = {0.5, 0.0, 0.0, -0.5, 0.0, 0.0};
    if (active) {
        if (HAS_PABS) {
            pabs = window[n] / dt;
            window[n] = 0.0;
        }
        te = (double)Te[n];
        tl = (double)Tl[n];
        tsT = (double)TsT[n];
        tsR = (double)TsR[n];
        m[0] = (double)mTx[n]; m[1] = (double)mTy[n]; m[2] = (double)mTz[n];
        m[3] = (double)mRx[n]; m[4] = (double)mRy[n]; m[5] = (double)mRz[n];
    }
    float f[MP_MLP_IN];
    cell_features(f, te, tl, tsT, tsR, m[0], m[3], pabs, dt, gd);
    stage_features(xs[w][l], f, norm);
    float r[MP_MLP_OUT];
    wave_infer(s, xs[w], ys[w], norm, r);  // contains the block's barriers
    if (!active) return;

    const double sub_dt = dt / (double)subcycles;

    // Check of the network's whole step against one plain substep taken at its midpoint (the
    // midpoint rule, second order like the step it checks): temperatures and m_x of the
    // midpoint state lie halfway along the predicted step, the transverse m at their start
    // value.  With r_mid the plain one-substep rate there, the predicted change must match
    // dt r_mid to tol (absolute in m, relative in T).  The transverse m, which the network
    // does not predict, advance by dt r_mid.
    bool accept = tol >= 0.0;
    const double hg = 0.5 * gd.gamma_e;
    const double ue0 = hg * (te * te);
    const double S = (double)subcycles;
    double e_te = 0, e_tl = 0, e_tsT = 0, e_tsR = 0, e_m[6] = {0, 0, 0, 0, 0, 0};
    if (accept) {
        e_te = sqrt(jl_max(hg, ue0 + dt * (pabs + (double)r[0])) / hg);
        e_tl = jl_max(1.0, tl + dt * (double)r[1]);
        e_tsT = jl_max(1.0, tsT + dt * (double)r[2]);
        e_tsR = jl_max(1.0, tsR + dt * (double)r[3]);
        e_m[0] = m[0] + dt * (double)r[4];
        e_m[3] = m[3] + dt * (double)r[5];
        accept = __builtin_isfinite(e_te) && __builtin_isfinite(e_tl) && __builtin_isfinite(e_tsT) &&
                 __builtin_isfinite(e_tsR) && __builtin_fabs(e_m[0]) <= 1.0 && __builtin_fabs(e_m[3]) <= 1.0;
    }
    if (accept) {
        const double h_te = 0.5 * (te + e_te), h_tl = 0.5 * (tl + e_tl), h_tsT = 0.5 * (tsT + e_tsT),
                     h_tsR = 0.5 * (tsR + e_tsR);
        const double h_m[6] = {0.5 * (m[0] + e_m[0]), m[1], m[2], 0.5 * (m[3] + e_m[3]), m[4], m[5]};
        double b_te = h_te, b_tl = h_tl, b_tsT = h_tsT, b_tsR = h_tsR;
        double b_m[6] = {h_m[0], h_m[1], h_m[2], h_m[3], h_m[4], h_m[5]};
        plain_substep<L>(b_te, b_tl, b_tsT, b_tsR, b_m, pabs, gd, tm_lut, re_lut, T_min, inv_dT, lut_N, sub_dt, iters);
        double res = __builtin_fabs(e_te - te - S * (b_te - h_te)) / e_te;
        res = jl_max(res, __builtin_fabs(e_tl - tl - S * (b_tl - h_tl)) / e_tl);
        res = jl_max(res, __builtin_fabs(e_tsT - tsT - S * (b_tsT - h_tsT)) / e_tsT);
        res = jl_max(res, __builtin_fabs(e_tsR - tsR - S * (b_tsR - h_tsR)) / e_tsR);
        res = jl_max(res, __builtin_fabs(e_m[0] - m[0] - S * (b_m[0] - h_m[0])));
        res = jl_max(res, __builtin_fabs(e_m[3] - m[3] - S * (b_m[3] - h_m[3])));
        e_m[1] = m[1] + S * (b_m[1] - h_m[1]);
        e_m[2] = m[2] + S * (b_m[2] - h_m[2]);
        e_m[4] = m[4] + S * (b_m[4] - h_m[4]);
        e_m[5] = m[5] + S * (b_m[5] - h_m[5]);
        accept = res <= tol;  // false for NaN
    }
    const unsigned long long mask = __ballot(!accept);
    if (mask) {
        // one list slot range per wave, lanes in order inside it
        const int leader = __ffsll((long long)__ballot(1)) - 1;
        uint32_t lo = 0, hi = 0;
        if (l == leader) {
            const unsigned long long base = atomicAdd(nrec, (unsigned long long)__popcll(mask));
            lo = (uint32_t)base;
            hi = (uint32_t)(base >> 32);
            if (fallbacks) atomicAdd(fallbacks, (unsigned long long)__popcll(mask));
        }
        lo = __shfl(lo, leader);
        hi = __shfl(hi, leader);
        if (!accept) {
            const unsigned long long slot = (((unsigned long long)hi << 32) | lo) +
                                            (unsigned long long)__popcll(mask & ((1ull << l) - 1ull));
            FallbackRec &rec = recs[slot];
            rec.s[0] = te; rec.s[1] = tl; rec.s[2] = tsT; rec.s[3] = tsR;
#pragma unroll
            for (int c = 0; c < 6; ++c) rec.s[4 + c] = m[c];
            rec.s[10] = pabs;
            rec.cell = n;
        }
    }
    if (accept) {
        Te[n] = (L)e_te; Tl[n] = (L)e_tl; TsT[n] = (L)e_tsT; TsR[n] = (L)e_tsR;
        mTx[n] = (L)e_m[0]; mTy[n] = (L)e_m[1]; mTz[n] = (L)e_m[2];
        mRx[n] = (L)e_m[3]; mRy[n] = (L)e_m[4]; mRz[n] = (L)e_m[5];
    }
    }
}

// ---- launch helpers --------------------------------------------------------------------

static float *g_train_part = nullptr;
static int64_t g_train_part_waves = 0;

static int check_mlp(const mp_mlp *m, const char *who) {
    if (!m || !m->params || !m->packed || !m->norm) return fail_arg("%s: incomplete mp_mlp", who);
    return MP_OK;
}

template <typename L>
static int launch_dataset(const mp_film *f, const mp_ferri_params *p, const mp_lut *lut, double dt, int subcycles,
                          int iters, int has_pabs, int64_t stride, float *X, float *T, int64_t rows) {
    const double dtc = (double)(L)dt, tmin = (double)(L)lut->T_min, invdt = (double)(L)lut->inv_dT;
    const FilmConsts *gc = film_consts(p);
    if (!gc) return MP_ERR_HIP;
    KernelTimer t("surrogate_dataset");
#define DS_ARGS (const L *)f->Te, (const L *)f->Tl, (const L *)f->Ts_TM, (const L *)f->Ts_RE, (const L *)f->m_TM_x, \
                (const L *)f->m_TM_y, (const L *)f->m_TM_z, (const L *)f->m_RE_x, (const L *)f->m_RE_y,         \
                (const L *)f->m_RE_z, f->window_sum, gc, (const L *)lut->tm, (const L *)lut->re, tmin, invdt,   \
                lut->n, dtc, subcycles, iters, stride, rows, X, T
    if (has_pabs) dataset_kernel<L, true><<<cdiv(rows, 128), 128, 0, stream()>>>(DS_ARGS);
    else dataset_kernel<L, false><<<cdiv(rows, 128), 128, 0, stream()>>>(DS_ARGS);
#undef DS_ARGS
    MP_LAUNCH_CHECK("dataset_kernel");
    return MP_OK;
}

static FallbackRec *g_recs = nullptr;
static int64_t g_recs_cap = 0;
static unsigned long long *g_nrec = nullptr;

static int ensure_fallback_list(int64_t n) {
    if (!g_nrec) MP_HIP(hipMalloc((void **)&g_nrec, sizeof(unsigned long long)));
    if (g_recs_cap < n) {
        if (g_recs) MP_HIP(hipFree(g_recs));
        g_recs = nullptr;
        g_recs_cap = 0;
        MP_HIP(hipMalloc((void **)&g_recs, sizeof(FallbackRec) * (size_t)n));
        g_recs_cap = n;
    }
    return MP_OK;
}

template <typename L>
static int launch_surrogate(mp_film *f, const mp_ferri_params *p, const mp_lut *lut, const mp_mlp *m, double dt,
                            int subcycles, int iters, int has_pabs, double tol, int64_t *fallbacks) {
    const int64_t N = f->n;
    const double dtc = (double)(L)dt, tmin = (double)(L)lut->T_min, invdt = (double)(L)lut->inv_dT;
    if (int e = ensure_fallback_list(N)) return e;
    const FilmConsts *gc = film_consts(p);
    if (!gc) return MP_ERR_HIP;
    reset_count_kernel<<<1, 1, 0, stream()>>>(g_nrec);
    MP_LAUNCH_CHECK("reset_count_kernel");
    KernelTimer t("surrogate");
#define SG_ARGS (L *)f->Te, (L *)f->Tl, (L *)f->Ts_TM, (L *)f->Ts_RE, (L *)f->m_TM_x, (L *)f->m_TM_y, \
                (L *)f->m_TM_z, (L *)f->m_RE_x, (L *)f->m_RE_y, (L *)f->m_RE_z, f->window_sum, gc,    \
                (const L *)lut->tm, (const L *)lut->re, tmin, invdt, lut->n, dtc, subcycles, iters, N, \
                (const uint2 *)m->packed, m->params, m->norm, tol, (unsigned long long *)fallbacks, g_recs, g_nrec
    // at most SUR_BLOCKS blocks resident for the whole update, looping over the cells: the
    // update then holds a fixed share of each SIMD beside the Yee kernels of the next step
    // instead of filling every CU (10 240 blocks at the 2.62 M-cell film)
    const unsigned grid = (unsigned)std::min<int64_t>(cdiv(N, 256), SUR_BLOCKS);
    if (has_pabs) surrogate_kernel<L, true><<<grid, 256, 0, stream()>>>(SG_ARGS);
    else surrogate_kernel<L, false><<<grid, 256, 0, stream()>>>(SG_ARGS);
#undef SG_ARGS
    MP_LAUNCH_CHECK("surrogate_kernel");
    {
        // the cells that failed the check take the plain integrator's substeps from their
        // start state; the grid loops over the list, so a step with few fallbacks costs a
        // small launch (it waits for room beside the next step's Yee kernels)
        KernelTimer t2("surrogate_fallback");
        const int64_t blocks = std::min<int64_t>(cdiv(N, 128), 512);
        surrogate_fallback_kernel<L><<<blocks, 128, 0, stream()>>>(
            (L *)f->Te, (L *)f->Tl, (L *)f->Ts_TM, (L *)f->Ts_RE, (L *)f->m_TM_x, (L *)f->m_TM_y, (L *)f->m_TM_z,
            (L *)f->m_RE_x, (L *)f->m_RE_y, (L *)f->m_RE_z, g_recs, g_nrec, gc, (const L *)lut->tm,
            (const L *)lut->re, tmin, invdt, lut->n, dtc / (double)subcycles, subcycles, iters);
        MP_LAUNCH_CHECK("surrogate_fallback_kernel");
    }
    return MP_OK;
}

static int check_film_step(const mp_film *f, const mp_ferri_params *p, const mp_lut *lut, int subcycles,
                           int has_pabs, const char *who) {
    if (!f || !p || !lut) return fail_arg("%s: null argument", who);
    if (subcycles <= 0) return fail_arg("%s: subcycles must be positive", who);
    if (f->n > 0) {
        if (lut->n < 1 || !lut->tm || !lut->re) return fail_arg("%s: empty LUT", who);
        if (lut->dtype != f->dtype) return fail_arg("%s: LUT dtype must equal film dtype", who);
        if (has_pabs && !f->window_sum) return fail_arg("%s: missing window_sum", who);
        if (f->dtype != MP_F32 && f->dtype != MP_F64) return fail_arg("%s: unsupported film dtype %d", who, f->dtype);
    }
    return MP_OK;
}

}  // namespace mp

using namespace mp;

MP_API int mp_mlp_pack(mp_mlp *m) {
    if (int e = check_mlp(m, "mp_mlp_pack")) return e;
    KernelTimer t("mlp_pack");
    mlp_pack_kernel<<<NFRAG, 64, 0, stream()>>>(m->params, (uint2 *)m->packed);
    MP_LAUNCH_CHECK("mlp_pack_kernel");
    return MP_OK;
}

MP_API int mp_mlp_forward(const mp_mlp *m, const float *X, float *Y, int64_t B) {
    if (int e = check_mlp(m, "mp_mlp_forward")) return e;
    if (B < 0) return fail_arg("mp_mlp_forward: negative batch");
    if (B == 0) return MP_OK;
    if (!X || !Y) return fail_arg("mp_mlp_forward: null array");
    KernelTimer t("mlp_forward");
    mlp_forward_kernel<<<cdiv(B, 256), 256, 0, stream()>>>((const uint2 *)m->packed, m->params, m->norm, X, Y, B);
    MP_LAUNCH_CHECK("mlp_forward_kernel");
    return MP_OK;
}

MP_API int mp_mlp_backward(const mp_mlp *m, const float *X, const float *T, int64_t B, float *grad, double *loss) {
    if (int e = check_mlp(m, "mp_mlp_backward")) return e;
    if (B <= 0) return fail_arg("mp_mlp_backward: batch must be positive");
    if (!X || !T || !grad) return fail_arg("mp_mlp_backward: null array");
    const int64_t tiles = cdiv(B, 16);
    const int64_t blocks = std::min<int64_t>(TRAIN_BLOCKS_MAX, cdiv(tiles, 4));
    const int64_t nwaves = blocks * 4;
    const int64_t per_wave = cdiv(tiles, nwaves);
    if (g_train_part_waves < nwaves) {
        if (g_train_part) MP_HIP(hipFree(g_train_part));
        g_train_part = nullptr;
        g_train_part_waves = 0;
        MP_HIP(hipMalloc((void **)&g_train_part, sizeof(float) * PART_STRIDE * TRAIN_BLOCKS_MAX * 4));
        g_train_part_waves = TRAIN_BLOCKS_MAX * 4;
    }
    const float gscale = (float)(2.0 / ((double)B * MP_MLP_OUT));
    {
        KernelTimer t("mlp_train");
        mlp_train_kernel<<<blocks, 256, 0, stream()>>>((const uint2 *)m->packed, m->params, m->norm, X, T, B, per_wave,
                                                       gscale, g_train_part);
        MP_LAUNCH_CHECK("mlp_train_kernel");
    }
    KernelTimer t("mlp_reduce");
    mlp_reduce_kernel<<<cdiv(PART_STRIDE, 256), 256, 0, stream()>>>(g_train_part, nwaves, grad, loss,
                                                                    1.0 / ((double)B * MP_MLP_OUT));
    MP_LAUNCH_CHECK("mlp_reduce_kernel");
    return MP_OK;
}

MP_API int mp_mlp_adam(float *params, const float *grad, float *m1, float *m2, int64_t step, float lr, float beta1,
                       float beta2, float eps, float wd) {
    if (!params || !grad || !m1 || !m2) return fail_arg("mp_mlp_adam: null array");
    if (step < 1) return fail_arg("mp_mlp_adam: step must be >= 1");
    const float c1 = (float)(1.0 - std::pow((double)beta1, (double)step));
    const float c2 = (float)(1.0 - std::pow((double)beta2, (double)step));
    KernelTimer t("mlp_adam");
    mlp_adam_kernel<<<cdiv(NPARAM, 256), 256, 0, stream()>>>(params, grad, m1, m2, lr, beta1, beta2, eps, wd, c1, c2);
    MP_LAUNCH_CHECK("mlp_adam_kernel");
    return MP_OK;
}

MP_API int mp_surrogate_dataset(const mp_film *f, const mp_ferri_params *p, const mp_lut *lut, double dt,
                                int32_t subcycles, int32_t brillouin_iters, int32_t has_pabs, int64_t stride,
                                float *X, float *T, int64_t max_rows) {
    if (int e = check_film_step(f, p, lut, subcycles, has_pabs, "mp_surrogate_dataset")) return e;
    if (stride < 1) return fail_arg("mp_surrogate_dataset: stride must be >= 1");
    if (max_rows < 0) return fail_arg("mp_surrogate_dataset: negative max_rows");
    const int64_t rows = std::min<int64_t>(cdiv(f->n, stride), max_rows);
    if (rows <= 0) return MP_OK;
    if (!X || !T) return fail_arg("mp_surrogate_dataset: null array");
    if (f->dtype == MP_F32)
        return launch_dataset<float>(f, p, lut, dt, subcycles, brillouin_iters, has_pabs, stride, X, T, rows);
    return launch_dataset<double>(f, p, lut, dt, subcycles, brillouin_iters, has_pabs, stride, X, T, rows);
}

MP_API int mp_surrogate_step(mp_film *f, const mp_ferri_params *p, const mp_lut *lut, const mp_mlp *m, double dt,
                             int32_t subcycles, int32_t brillouin_iters, int32_t has_pabs, double tol,
                             int64_t *fallbacks) {
    if (int e = check_film_step(f, p, lut, subcycles, has_pabs, "mp_surrogate_step")) return e;
    if (int e = check_mlp(m, "mp_surrogate_step")) return e;
    if (f->n <= 0) return MP_OK;
    if (f->dtype == MP_F32)
        return launch_surrogate<float>(f, p, lut, m, dt, subcycles, brillouin_iters, has_pabs, tol, fallbacks);
    return launch_surrogate<double>(f, p, lut, m, dt, subcycles, brillouin_iters, has_pabs, tol, fallbacks);
}
