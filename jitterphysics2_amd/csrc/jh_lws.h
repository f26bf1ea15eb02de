// Learned warm start: a per-contact-point MLP 24 -> 128 -> 128 -> 3 (GELU) that predicts the
// accumulated impulse (normal, tangent1, tangent2) the solver converges to, used in place of the
// previous-frame impulse when a step starts.  BF16 weights and activations, FP32 accumulation, on
// v_mfma_f32_16x16x16_bf16.  The reference has no counterpart (it warm-starts from the previous
// frame, Contact.cs:802-817); the plain warm start stays the fallback.
//
// MFMA formulation.  Every layer is computed transposed, Z^T[out][sample] = W[out][in] . H^T[in][sample],
// on 16 x 16 tiles with 16 samples per wave.  For v_mfma_f32_16x16x16_bf16, lane l holds
//   A[m = l % 16][k = 4 (l / 16) + r],  B[k = 4 (l / 16) + r][n = l % 16],  D[m = 4 (l / 16) + r][n = l % 16]
// for r = 0..3, so the D registers of output tile t are exactly the B registers of K-step t of the
// next layer: activations never leave registers between layers.  Weights are read as A operands
// from global memory in PyTorch Linear layout W[out][in] (forward) and transposed copies W^T[in][out]
// (backward), all BF16.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/jitterhip.h"
#include "jh_kernels.h"
#include "jh_math.h"

namespace jh {
namespace lws {

constexpr int kIn = 24, kInPad = 32, kHid = 128, kOut = 3, kOutPad = 16;
// flat FP32 parameter vector, PyTorch Linear order: W1[128][24], b1[128], W2[128][128], b2[128],
// W3[3][128], b3[3]
constexpr int oW1 = 0, oB1 = oW1 + kHid * kIn, oW2 = oB1 + kHid, oB2 = oW2 + kHid * kHid, oW3 = oB2 + kHid,
              oB3 = oW3 + kOut * kHid, kParams = oB3 + kOut;
static_assert(kParams == 20099, "parameter count");
constexpr int kBiasCols = kHid + kHid + kOutPad;  // per-wave bias-gradient partials: db1 | db2 | db3

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (NaN -> quiet NaN).
__host__ __device__ __forceinline__ uint16_t f2bf(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    uint32_t u = __float_as_uint(x);
#else
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
#endif
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x4 ld4(const uint16_t* p) { return *reinterpret_cast<const bf16x4*>(p); }
__device__ __forceinline__ bf16x4 pack4(f32x4 v) {
    bf16x4 r;
    r[0] = short(f2bf(v[0]));
    r[1] = short(f2bf(v[1]));
    r[2] = short(f2bf(v[2]));
    r[3] = short(f2bf(v[3]));
    return r;
}

// GELU(x) = x * Phi(x) and its derivative Phi(x) + x * phi(x).
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// Device copy of a model: BF16 weights in both layouts, FP32 biases, feature normalisation.
struct Model {
    const uint16_t* w1;   // [128][32]   W1, inputs padded with zeros
    const uint16_t* w2;   // [128][128]
    const uint16_t* w3;   // [16][128]   rows 3..15 zero
    const uint16_t* w2t;  // [128][128]  W2^T
    const uint16_t* w3t;  // [128][16]   W3^T, columns 3..15 zero
    const float* b1;      // [128]
    const float* b2;      // [128]
    const float* b3;      // [16]        entries 3..15 zero
    const float* mean;    // [24] feature mean
    const float* istd;    // [24] 1 / feature standard deviation
    const float* tscale;  // [3]  output scale: impulse = previous + y * tscale
};

struct Fwd {
    f32x4 z1[8], z2[8];
    bf16x4 h1[8], h2[8];
    f32x4 y;
};

// Forward pass of 16 samples; x[kt] = B operand of input K-step kt (features 16 kt .. 16 kt + 15).
__device__ __forceinline__ void forward16(const Model& M, const bf16x4 x[2], Fwd& F, int lane) {
    int row = lane & 15, kq = (lane >> 4) * 4;
#pragma unroll
    for (int mt = 0; mt < 8; ++mt) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kt = 0; kt < 2; ++kt) acc = mfma16(ld4(M.w1 + (mt * 16 + row) * kInPad + kt * 16 + kq), x[kt], acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r] += M.b1[mt * 16 + kq + r];
        F.z1[mt] = acc;
        f32x4 h;
#pragma unroll
        for (int r = 0; r < 4; ++r) h[r] = gelu(acc[r]);
        F.h1[mt] = pack4(h);
    }
#pragma unroll
    for (int mt = 0; mt < 8; ++mt) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kt = 0; kt < 8; ++kt) acc = mfma16(ld4(M.w2 + (mt * 16 + row) * kHid + kt * 16 + kq), F.h1[kt], acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r] += M.b2[mt * 16 + kq + r];
        F.z2[mt] = acc;
        f32x4 h;
#pragma unroll
        for (int r = 0; r < 4; ++r) h[r] = gelu(acc[r]);
        F.h2[mt] = pack4(h);
    }
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kt = 0; kt < 8; ++kt) acc = mfma16(ld4(M.w3 + row * kHid + kt * 16 + kq), F.h2[kt], acc);
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] += M.b3[kq + r];
    F.y = acc;  // lanes 0..15: outputs 0..2 of sample `lane` in y[0..2]
}

// ---- Contact features -----------------------------------------------------------------------
// 24 features of contact slot s at the warm start of the first substep (velocities after
// IntegrateForces, World.Step.cs:1060-1072):
//   0-2 normal, 3-5 tangent1, 6-8 tangent2, 9-11 relative contact velocity (dv . n, dv . t1, dv . t2),
//   12 penetration (as PrepareForIteration computes it, Contact.cs:831-840), 13-15 MassNormalTangent,
//   16-18 Accumulated (previous frame), 19 substep index, 20 friction, 21 Bias, 22 PenaltyBias,
//   23 NewContact flag.
template <class BodiesT>
__device__ __forceinline__ void contact_features(const jh_arbiter& a, int s, const BodiesT& B, float f[kIn]) {
    const jh_contact& c = a.contacts[s];
    int b1 = a.body1, b2 = a.body2;
    vec3 v1 = ld3(B.vel[b1]), w1 = ld3(B.angvel[b1]);
    vec3 v2 = ld3(B.vel[b2]), w2 = ld3(B.angvel[b2]);
    if (B.motion[b1] == JH_MOTION_DYNAMIC) { w1 = w1 + ld3(B.dangvel[b1]); v1 = v1 + ld3(B.dvel[b1]); }
    if (B.motion[b2] == JH_MOTION_DYNAMIC) { w2 = w2 + ld3(B.dangvel[b2]); v2 = v2 + ld3(B.dvel[b2]); }
    vec3 n{c.normal[0], c.normal[1], c.normal[2]};
    vec3 t1{c.tangent1[0], c.tangent1[1], c.tangent1[2]};
    vec3 t2{c.tangent2[0], c.tangent2[1], c.tangent2[2]};
    vec3 rel1{c.relative_position1[0], c.relative_position1[1], c.relative_position1[2]};
    vec3 rel2{c.relative_position2[0], c.relative_position2[1], c.relative_position2[2]};
    vec3 dv = v2 + cross(w2, rel2);
    dv = dv - (v1 + cross(w1, rel1));
    vec3 r1 = rot(vec3{c.position1[0], c.position1[1], c.position1[2]}, ldq(B.quat[b1])) + ld3(B.pos[b1]);
    vec3 r2 = rot(vec3{c.position2[0], c.position2[1], c.position2[2]}, ldq(B.quat[b2])) + ld3(B.pos[b2]);
    f[0] = n.x; f[1] = n.y; f[2] = n.z;
    f[3] = t1.x; f[4] = t1.y; f[5] = t1.z;
    f[6] = t2.x; f[7] = t2.y; f[8] = t2.z;
    f[9] = dot(dv, n); f[10] = dot(dv, t1); f[11] = dot(dv, t2);
    f[12] = dot(r1 - r2, n);
    f[13] = c.mass_normal_tangent[0]; f[14] = c.mass_normal_tangent[1]; f[15] = c.mass_normal_tangent[2];
    f[16] = c.accumulated[0]; f[17] = c.accumulated[1]; f[18] = c.accumulated[2];
    f[19] = 0.0f;
    f[20] = a.friction;
    f[21] = c.bias;
    f[22] = c.penalty_bias;
    f[23] = (c.flag & 2u) ? 1.0f : 0.0f;
}

// Per-contact impulse scale of the network output: effective normal mass x (relative speed +
// penalty bias + 0.05 m/s), from the raw features.  Zero for contacts without an effective mass yet
// (new contacts), which therefore keep the previous-frame warm start.
__host__ __device__ __forceinline__ float impulse_scale(const float* f) {
    return fabsf(f[13]) * (fabsf(f[9]) + fabsf(f[10]) + fabsf(f[11]) + fabsf(f[22]) + 0.05f);
}

// Whether the solver visits the arbiter: a body of it is non-static to the solver (B.motion holds
// the solver's motion types, a sleeping body static).
template <class BodiesT>
__device__ __forceinline__ bool arbiter_solved(const jh_arbiter& a, const BodiesT& B) {
    return B.motion[a.body1] != JH_MOTION_STATIC || B.motion[a.body2] != JH_MOTION_STATIC;
}

// Raw features of every solved slot (row = 4 * arbiter + slot); valid[row] = slot in use (other rows
// of `out` are left untouched).
template <class BodiesT>
__global__ void __launch_bounds__(256) k_features(const jh_arbiter* __restrict__ arb, const int* __restrict__ a_count,
                                                  BodiesT B, float* __restrict__ out, uint8_t* __restrict__ valid) {
    int n = *a_count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 4 * n; i += gridDim.x * blockDim.x) {
        const jh_arbiter& a = arb[i >> 2];
        int s = i & 3;
        bool used = (a.usage_mask & (1u << s)) != 0 && arbiter_solved(a, B);
        valid[i] = used ? 1 : 0;
        if (!used) continue;
        float f[kIn];
        contact_features(a, s, B, f);
        for (int k = 0; k < kIn; ++k) out[size_t(i) * kIn + k] = f[k];
    }
}

// Accumulated impulses of every slot (row = 4 * arbiter + slot).
__global__ void __launch_bounds__(256) k_gather_acc(const jh_arbiter* __restrict__ arb, const int* __restrict__ a_count,
                                                    float* __restrict__ out) {
    int n = *a_count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 4 * n; i += gridDim.x * blockDim.x) {
        const jh_contact& c = arb[i >> 2].contacts[i & 3];
        out[3 * size_t(i) + 0] = c.accumulated[0];
        out[3 * size_t(i) + 1] = c.accumulated[1];
        out[3 * size_t(i) + 2] = c.accumulated[2];
    }
}

__global__ void __launch_bounds__(256) k_zero_acc(jh_arbiter* __restrict__ arb, const int* __restrict__ a_count) {
    int n = *a_count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 4 * n; i += gridDim.x * blockDim.x) {
        jh_contact& c = arb[i >> 2].contacts[i & 3];
        c.accumulated[0] = 0.0f;
        c.accumulated[1] = 0.0f;
        c.accumulated[2] = 0.0f;
    }
}

// Warm start from the network: each wave takes 64 slots, every lane builds the normalised BF16
// features of one slot in LDS, then four 16-sample MLP passes; the prediction replaces Accumulated
// for used slots that will be warm-started (not SkipWarmStart).  prediction = previous + y * tscale
// * impulse_scale with y limited to +-kTrust standard deviations of the training corrections (a trust region around
// the previous-frame impulse for states unlike the training data), normal impulse clamped at 0; a
// non-finite prediction keeps the previous impulse.
constexpr float kTrust = 3.0f;
template <class BodiesT>
__global__ void __launch_bounds__(256) k_infer(jh_arbiter* __restrict__ arb, const int* __restrict__ a_count, BodiesT B,
                                               Model M) {
    __shared__ uint16_t xs[4][64][kInPad];
    __shared__ float ys[4][64][4];
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int n4 = 4 * *a_count;
    int nwaves = gridDim.x * 4;
    for (int base = (blockIdx.x * 4 + wv) * 64; base < n4; base += nwaves * 64) {
        int i = base + lane;
        bool use = false;
        float sc = 0.0f;
        if (i < n4) {
            const jh_arbiter& a = arb[i >> 2];
            int s = i & 3;
            use = (a.usage_mask & (1u << s)) && !(a.contacts[s].flag & 4u) && arbiter_solved(a, B);
            float f[kIn];
            if (use) {
                contact_features(a, s, B, f);
                sc = impulse_scale(f);
            }
            for (int k = 0; k < kInPad; ++k)
                xs[wv][lane][k] = (use && k < kIn) ? f2bf((f[k] - M.mean[k]) * M.istd[k]) : uint16_t(0);
        } else {
            for (int k = 0; k < kInPad; ++k) xs[wv][lane][k] = 0;
        }
        __builtin_amdgcn_wave_barrier();
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        for (int g = 0; g < 4; ++g) {
            int row = lane & 15, kq = (lane >> 4) * 4;
            bf16x4 x[2];
            x[0] = ld4(&xs[wv][g * 16 + row][kq]);
            x[1] = ld4(&xs[wv][g * 16 + row][16 + kq]);
            Fwd F;
            forward16(M, x, F, lane);
            if (lane < 16)
                for (int r = 0; r < 3; ++r) ys[wv][g * 16 + lane][r] = F.y[r];
        }
        __builtin_amdgcn_wave_barrier();
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if (use) {
            jh_contact& c = arb[i >> 2].contacts[i & 3];
            float p[3];
            bool ok = true;
            for (int r = 0; r < 3; ++r) {
                float y = fminf(fmaxf(ys[wv][lane][r], -kTrust), kTrust);
                p[r] = c.accumulated[r] + y * (M.tscale[r] * sc);
                ok = ok && (p[r] - p[r] == 0.0f);
            }
            if (ok) {
                c.accumulated[0] = fmaxf(p[0], 0.0f);
                c.accumulated[1] = p[1];
                c.accumulated[2] = p[2];
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ---- Training -------------------------------------------------------------------------------
// Training data on the device: xb[n][32] normalised BF16 features, tn[n][3] normalised targets
// (accumulated impulse after the reference iterations minus the previous-frame impulse, / tscale).
struct Batch {
    const uint16_t* xb;
    const float* tn;
    int begin, count;  // samples [begin, begin + count)
    int ld;            // row length of the feature-major buffers below (>= count, multiple of 16)
    float gscale;      // dLoss/dY = gscale * (Y - T): 2 / (3 count)
    // outputs, feature-major [F][ld] BF16 (operands of the weight-gradient GEMMs)
    uint16_t *xt, *h1t, *h2t, *dz1t, *dz2t, *dyt;
    float* bias_part;  // [waves][kBiasCols]
    double* loss_part; // [waves] sum of squared errors
};

__device__ __forceinline__ float red16(float v) {
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    v += __shfl_xor(v, 8, 64);
    return v;
}

// One wave per 16 samples: forward, loss, backward through the activations; writes the BF16
// operands of the weight gradients and per-wave bias-gradient and loss partials.
__global__ void __launch_bounds__(256) k_fwd_bwd(Model M, Batch T) {
    int lane = threadIdx.x & 63;
    int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    int s0 = w * 16;
    if (s0 >= T.ld) return;
    int row = lane & 15, kq = (lane >> 4) * 4;
    int sl = s0 + row;               // this lane's sample (local to the batch)
    bool valid = sl < T.count;
    size_t sg = size_t(T.begin) + size_t(valid ? sl : 0);
    bf16x4 x[2];
    for (int kt = 0; kt < 2; ++kt) {
        x[kt] = valid ? ld4(T.xb + sg * kInPad + kt * 16 + kq) : bf16x4{0, 0, 0, 0};
    }
    Fwd F;
    forward16(M, x, F, lane);
    // dLoss/dY on lanes 0..15 (outputs 0..2), zero elsewhere
    f32x4 dy = {0.0f, 0.0f, 0.0f, 0.0f};
    float sq = 0.0f;
    if (valid && kq == 0)
        for (int r = 0; r < 3; ++r) {
            float e = F.y[r] - T.tn[sg * 3 + r];
            sq += e * e;
            dy[r] = T.gscale * e;
        }
    bf16x4 dyb = pack4(dy);
    // dH2^T = W3^T . dY^T, dZ2 = dH2 * gelu'(Z2)
    bf16x4 dz2b[8];
    f32x4 dz2[8];
#pragma unroll
    for (int mt = 0; mt < 8; ++mt) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mfma16(ld4(M.w3t + (mt * 16 + row) * kOutPad + kq), dyb, acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r] *= gelu_grad(F.z2[mt][r]);
        dz2[mt] = acc;
        dz2b[mt] = pack4(acc);
    }
    // dH1^T = W2^T . dZ2^T, dZ1 = dH1 * gelu'(Z1)
    f32x4 dz1[8];
#pragma unroll
    for (int mt = 0; mt < 8; ++mt) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kt = 0; kt < 8; ++kt) acc = mfma16(ld4(M.w2t + (mt * 16 + row) * kHid + kt * 16 + kq), dz2b[kt], acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r] *= gelu_grad(F.z1[mt][r]);
        dz1[mt] = acc;
    }
    // feature-major stores: element (feature f, sample s0 + row)
    size_t col = size_t(s0 + row);
    auto st4 = [&](uint16_t* base, int f0, bf16x4 v) {
        for (int r = 0; r < 4; ++r) base[size_t(f0 + kq + r) * T.ld + col] = uint16_t(v[r]);
    };
    st4(T.xt, 0, x[0]);
    st4(T.xt, 16, x[1]);
    for (int mt = 0; mt < 8; ++mt) {
        st4(T.h1t, mt * 16, F.h1[mt]);
        st4(T.h2t, mt * 16, F.h2[mt]);
        st4(T.dz1t, mt * 16, pack4(dz1[mt]));
        st4(T.dz2t, mt * 16, dz2b[mt]);
    }
    st4(T.dyt, 0, dyb);
    // bias-gradient partials (FP32, before BF16 rounding) and loss
    float* bp = T.bias_part + size_t(w) * kBiasCols;
    for (int mt = 0; mt < 8; ++mt)
        for (int r = 0; r < 4; ++r) {
            float a = red16(dz1[mt][r]);
            float b = red16(dz2[mt][r]);
            if (row == 0) {
                bp[mt * 16 + kq + r] = a;
                bp[kHid + mt * 16 + kq + r] = b;
            }
        }
    for (int r = 0; r < 4; ++r) {
        float d = red16(dy[r]);
        if (row == 0) bp[2 * kHid + kq + r] = d;
    }
    float l = sq;
    for (int o = 1; o < 64; o <<= 1) l += __shfl_xor(l, o, 64);
    if (lane == 0) T.loss_part[w] = double(l);
}

// Split-K weight-gradient GEMM: C[m][n] = sum_k A[m][k] B[n][k] over k in this split, A and B
// feature-major BF16 with row length ld.  One wave per (16 x 16 tile, split); partial results go to
// part[split][M][N] and are summed in split order by the optimizer.
struct GemmJob {
    const uint16_t* A;
    const uint16_t* B;
    float* part;
    int M, N;
};
struct GemmArgs {
    GemmJob job[3];
    int tiles[3];   // (M / 16) * (N / 16) per job
    int splits, ksplit, ld;
};
__global__ void __launch_bounds__(256) k_grad_gemm(GemmArgs G) {
    int lane = threadIdx.x & 63;
    int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    int total = (G.tiles[0] + G.tiles[1] + G.tiles[2]) * G.splits;
    if (w >= total) return;
    int split = w % G.splits;
    int t = w / G.splits;
    int j = 0;
    while (t >= G.tiles[j]) { t -= G.tiles[j]; ++j; }
    const GemmJob& J = G.job[j];
    int nt = J.N / 16;
    int mt = t / nt, ntile = t % nt;
    int row = lane & 15, kq = (lane >> 4) * 4;
    const uint16_t* a = J.A + size_t(mt * 16 + row) * G.ld;
    const uint16_t* b = J.B + size_t(ntile * 16 + row) * G.ld;
    int k0 = split * G.ksplit, k1 = min(k0 + G.ksplit, G.ld);
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k = k0; k < k1; k += 16) acc = mfma16(ld4(a + k + kq), ld4(b + k + kq), acc);
    float* out = J.part + size_t(split) * J.M * J.N;
    for (int r = 0; r < 4; ++r) out[size_t(mt * 16 + kq + r) * J.N + ntile * 16 + row] = acc[r];
}

// Gradient of every parameter from the partials (fixed summation order), then optionally Adam.
struct OptArgs {
    float* params;           // [kParams] FP32 master copy
    float* grad;             // [kParams] output gradient
    float *m, *v;            // Adam moments
    const float* gw1;        // [splits][128][32]
    const float* gw2;        // [splits][128][128]
    const float* gw3;        // [splits][16][128]
    const float* bias_part;  // [waves][kBiasCols]
    int splits, waves;
    int adam;                // 0: gradient only
    float lr, beta1, beta2, eps, bc1, bc2;  // bc = 1 - beta^t
};
__global__ void __launch_bounds__(256) k_optimizer(OptArgs O) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kParams) return;
    float g = 0.0f;
    if (p < oB1) {
        int o = (p - oW1) / kIn, i = (p - oW1) % kIn;
        for (int s = 0; s < O.splits; ++s) g += O.gw1[(size_t(s) * kHid + o) * kInPad + i];
    } else if (p < oW2) {
        int o = p - oB1;
        for (int w = 0; w < O.waves; ++w) g += O.bias_part[size_t(w) * kBiasCols + o];
    } else if (p < oB2) {
        int o = (p - oW2) / kHid, i = (p - oW2) % kHid;
        for (int s = 0; s < O.splits; ++s) g += O.gw2[(size_t(s) * kHid + o) * kHid + i];
    } else if (p < oW3) {
        int o = p - oB2;
        for (int w = 0; w < O.waves; ++w) g += O.bias_part[size_t(w) * kBiasCols + kHid + o];
    } else if (p < oB3) {
        int o = (p - oW3) / kHid, i = (p - oW3) % kHid;
        for (int s = 0; s < O.splits; ++s) g += O.gw3[(size_t(s) * kOutPad + o) * kHid + i];
    } else {
        int o = p - oB3;
        for (int w = 0; w < O.waves; ++w) g += O.bias_part[size_t(w) * kBiasCols + 2 * kHid + o];
    }
    O.grad[p] = g;
    if (O.adam) {
        float m = O.beta1 * O.m[p] + (1.0f - O.beta1) * g;
        float v = O.beta2 * O.v[p] + (1.0f - O.beta2) * g * g;
        O.m[p] = m;
        O.v[p] = v;
        float mh = m / O.bc1, vh = v / O.bc2;
        O.params[p] -= O.lr * mh / (sqrtf(vh) + O.eps);
    }
}

// BF16 weight copies (both layouts) and padded FP32 biases from the FP32 parameters.
struct CopyArgs {
    const float* params;
    uint16_t *w1, *w2, *w3, *w2t, *w3t;
    float *b1, *b2, *b3;
};
__global__ void __launch_bounds__(256) k_refresh(CopyArgs C) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    const float* P = C.params;
    if (t < kHid * kInPad) {  // w1 [128][32]
        int o = t / kInPad, i = t % kInPad;
        C.w1[t] = i < kIn ? f2bf(P[oW1 + o * kIn + i]) : uint16_t(0);
    }
    if (t < kHid * kHid) {  // w2 and w2t
        int o = t / kHid, i = t % kHid;
        uint16_t h = f2bf(P[oW2 + t]);
        C.w2[t] = h;
        C.w2t[i * kHid + o] = h;
    }
    if (t < kOutPad * kHid) {  // w3 [16][128], w3t [128][16]
        int o = t / kHid, i = t % kHid;
        uint16_t h = o < kOut ? f2bf(P[oW3 + o * kHid + i]) : uint16_t(0);
        C.w3[t] = h;
        C.w3t[i * kOutPad + o] = h;
    }
    if (t < kHid) {
        C.b1[t] = P[oB1 + t];
        C.b2[t] = P[oB2 + t];
    }
    if (t < kOutPad) C.b3[t] = t < kOut ? P[oB3 + t] : 0.0f;
}

// Normalised training tensors from raw features [n][24] and raw targets [n][3] (accumulated impulse
// after the reference iterations): xb = bf16((f - mean) * istd),
// tn = (target - f[16..18]) / (tscale * impulse_scale(f)) (0 where the scale is 0).
__global__ void __launch_bounds__(256) k_prepare_data(const float* __restrict__ f, const float* __restrict__ t, int n,
                                                      const float* __restrict__ mean, const float* __restrict__ istd,
                                                      const float* __restrict__ tscale, uint16_t* __restrict__ xb,
                                                      float* __restrict__ tn) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < kInPad; ++k)
        xb[size_t(i) * kInPad + k] = k < kIn ? f2bf((f[size_t(i) * kIn + k] - mean[k]) * istd[k]) : uint16_t(0);
    if (t) {
        float sc = impulse_scale(f + size_t(i) * kIn);
        for (int r = 0; r < 3; ++r) {
            float d = t[size_t(i) * 3 + r] - f[size_t(i) * kIn + 16 + r];
            float den = tscale[r] * sc;
            tn[size_t(i) * 3 + r] = den > 0.0f ? d / den : 0.0f;
        }
    }
}

// Forward only (predictions in the normalised output space) for evaluation.
__global__ void __launch_bounds__(256) k_predict(Model M, const uint16_t* __restrict__ xb, int n,
                                                 float* __restrict__ y) {
    int lane = threadIdx.x & 63;
    int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    int s0 = w * 16;
    if (s0 >= n) return;
    int row = lane & 15, kq = (lane >> 4) * 4;
    int s = s0 + row;
    bool valid = s < n;
    bf16x4 x[2];
    for (int kt = 0; kt < 2; ++kt) x[kt] = valid ? ld4(xb + size_t(s) * kInPad + kt * 16 + kq) : bf16x4{0, 0, 0, 0};
    Fwd F;
    forward16(M, x, F, lane);
    if (valid && kq == 0)
        for (int r = 0; r < 3; ++r) y[size_t(s) * 3 + r] = F.y[r];
}

// ---- Residual of the velocity problem ---------------------------------------------------------
// Per contact point, with the bias SolveVelocities uses (max(PenaltyBias, Bias) when PenaltyBias > 0):
// e_n = bias - v_n if lambda_n > 0 else max(0, bias - v_n); e_t = |v_t| if |lambda_t| < friction *
// lambda_n(old) else 0, for both tangents, where lambda_n(old) is the normal impulse the friction
// clamp of the last iteration used (the value before that iteration; the current one when lam_old is
// NULL).  Returns the sum of squares.
template <class BodiesT>
__device__ __forceinline__ double arbiter_residual(const jh_arbiter& a, const BodiesT& B, const float* lam_old) {
    vec3 v1 = ld3(B.vel[a.body1]), w1 = ld3(B.angvel[a.body1]);
    vec3 v2 = ld3(B.vel[a.body2]), w2 = ld3(B.angvel[a.body2]);
    double s = 0.0;
    for (int k = 0; k < 4; ++k) {
        if (!(a.usage_mask & (1u << k))) continue;
        const jh_contact& c = a.contacts[k];
        vec3 rel1{c.relative_position1[0], c.relative_position1[1], c.relative_position1[2]};
        vec3 rel2{c.relative_position2[0], c.relative_position2[1], c.relative_position2[2]};
        vec3 dv = v2 + cross(w2, rel2);
        dv = dv - (v1 + cross(w1, rel1));
        float vn = c.normal[0] * dv.x + c.normal[1] * dv.y + c.normal[2] * dv.z;
        float vt1 = c.tangent1[0] * dv.x + c.tangent1[1] * dv.y + c.tangent1[2] * dv.z;
        float vt2 = c.tangent2[0] * dv.x + c.tangent2[1] * dv.y + c.tangent2[2] * dv.z;
        float bias = c.penalty_bias > 0.0f ? fmaxf(c.penalty_bias, c.bias) : c.bias;
        float ln = c.accumulated[0];
        float en = ln > 0.0f ? bias - vn : fmaxf(0.0f, bias - vn);
        float lim = a.friction * (lam_old ? lam_old[k] : ln);
        float et1 = fabsf(c.accumulated[1]) < lim ? fabsf(vt1) : 0.0f;
        float et2 = fabsf(c.accumulated[2]) < lim ? fabsf(vt2) : 0.0f;
        s += double(en) * en + double(et1) * et1 + double(et2) * et2;
    }
    return s;
}

// Residual trace of the first substep on a scratch copy of the velocities and arbiters:
// IntegrateForces, PrepareForIteration with whatever Accumulated the copy holds (the warm start
// under test), then K iterations of SolveVelocities in the colour batches of k_solve, recording the
// residual sum of squares after the prepare (k = 0) and after each iteration: part[k][block].
// lam_old[4 * arbiter + slot] keeps the normal impulses from before the current iteration.
__global__ void __launch_bounds__(256) k_probe(SolveArgs P, int K, double* __restrict__ part, float* __restrict__ lam_old) {
    __shared__ double red[256];
    int stride = gridDim.x * blockDim.x;
    int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    int nc = *P.ncolours;
    int na = P.offs[kMaxColours + 1];
    integrate_forces(P, t0, stride);
    if (!grid_barrier(P.bar, P.spin)) return;
    for (int k = 0; k <= K; ++k) {
        if (k > 0) {
            for (int i = t0; i < na; i += stride) {
                uint32_t ai = uint32_t(P.order[i]);
                for (int s = 0; s < 4; ++s) lam_old[4 * size_t(ai) + s] = P.arb[ai].contacts[s].accumulated[0];
            }
            if (!grid_barrier(P.bar, P.spin)) return;
        }
        if (!solve_batches(P, k == 0 ? 0 : 1, nc)) return;
        double s = 0.0;
        for (int i = t0; i < na; i += stride) {
            uint32_t ai = uint32_t(P.order[i]);
            s += arbiter_residual(P.arb[ai], P.B, k > 0 ? lam_old + 4 * size_t(ai) : nullptr);
        }
        red[threadIdx.x] = s;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[size_t(k) * gridDim.x + blockIdx.x] = red[0];
        if (!grid_barrier(P.bar, P.spin)) return;
    }
}

__global__ void k_sum_loss(const double* __restrict__ part, int n, double* __restrict__ acc) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += part[i];
        *acc += s;
    }
}

}  // namespace lws
}  // namespace jh
