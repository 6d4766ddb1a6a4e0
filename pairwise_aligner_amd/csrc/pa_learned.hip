// pa_learned.hip -- learned lane dispatch: per-pair features and a 3-layer FP32 MLP with its own
// forward, backward and Adam kernels (no library GEMM: the network is 8 -> 64 -> 64 -> 2 and one
// thread carries one sample through it, with the weights staged in LDS).
//
// Features of pair (s1, s2) with lengths (m, n), computed by one wave per pair:
//   0 log2(m + 1) / 16           1 log2(n + 1) / 16          2 min(m, n) / max(m, n)
//   3 hashed 3-mer composition distance: L1 distance of the normalised 64-bin histograms / 2
//   4 minimizer containment: the fraction of s2's (w = 8, k)-minimizers -- the smallest hashed
//     k-mer of every window of w consecutive k-mers -- found in a 32768-bit set of s1's
//     (k = 8 over 2-bit codes for alphabets of at most four symbols, else k = 3 over 5-bit codes)
//   5 identity estimate containment^(1/k)
//   6 max substitution score / 16                        7 1
// Outputs: y0 = max of H over the DP / (max substitution score * min(m, n)), y1 = log2(cost)/32.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pa_learned.hpp"

namespace pa {

constexpr int F = kModelFeatures, HID = kModelHidden, OUT = kModelOutputs;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + F * HID, OFF_W2 = OFF_B1 + HID, OFF_B2 = OFF_W2 + HID * HID,
              OFF_W3 = OFF_B2 + HID, OFF_B3 = OFF_W3 + HID * OUT, NPARAM = OFF_B3 + OUT;
static_assert(NPARAM == kModelParams, "parameter layout");

constexpr int kSetWords = 1024;   // 32768-bit minimizer set per wave
constexpr int kMinWindow = 8;     // k-mers per minimizer window

__device__ __forceinline__ uint32_t mix32(uint32_t x)
{
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}

// The (kMinWindow, k)-minimizers of sequence s (length len), visited by one wave: for every
// minimizer, in order, visit(hash) is called on one lane.  A tile of 64 consecutive k-mers is
// loaded one per lane; window j (its first k-mer at tile position j <= 64 - kMinWindow) takes the
// smallest (hash, position) key of its k-mers by three shuffle steps; a window whose minimizer
// sits at another position than its predecessor's contributes a new minimizer (the positions of
// sliding-window minimizers never decrease).  Ties go to the leftmost k-mer.  A sequence with
// fewer k-mers than a window has one window over all of them.
template <typename Visit>
__device__ __forceinline__ void for_each_minimizer(const uint8_t *s, int len, const FeatureArgs &a, Visit visit)
{
    const int lane = threadIdx.x % 64;
    const int k = a.k, bits = a.bits;
    const uint32_t kmask = (1u << (k * bits)) - 1u, smask = (1u << bits) - 1u;
    const int kmers = len - k + 1;
    if (kmers <= 0) return;
    const int windows = kmers > kMinWindow ? kmers - kMinWindow + 1 : 1;
    constexpr int kStep = 64 - kMinWindow + 1;   // windows per tile
    uint64_t prev = ~0ull;                        // key of the previous window's minimizer
    for (int base = 0; base < windows; base += kStep) {
        const int i = base + lane;                // this lane's k-mer
        uint64_t key = ~0ull;
        if (i < kmers) {
            uint32_t v = 0;
            for (int j = 0; j < k; ++j) v = (v << bits) | (uint32_t(a.code[s[i + j]]) & smask);
            key = (uint64_t(mix32(v & kmask)) << 32) | uint32_t(i);
        }
        // minimum over the kMinWindow k-mers starting at this lane
        for (int d = 1; d < kMinWindow; d <<= 1) {
            const uint64_t o = __shfl_down(key, d, 64);
            key = (lane + d < 64 && o < key) ? o : key;
        }
        const bool valid = lane < kStep && base + lane < windows;
        uint64_t before = __shfl_up(key, 1, 64);
        if (lane == 0) before = prev;
        if (valid && key != before) visit(uint32_t(key >> 32));
        prev = __shfl(key, kStep - 1, 64);
    }
}

__global__ __launch_bounds__(256) void features_kernel(FeatureArgs a)
{
    __shared__ uint32_t comp[4][2][64];
    __shared__ uint32_t kset[4][kSetWords];
    __shared__ uint32_t hits[4], counts[4];
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    const uint64_t p = uint64_t(blockIdx.x) * 4 + w;
    for (int i = lane; i < 128; i += 64) comp[w][i / 64][i % 64] = 0;
    for (int i = lane; i < kSetWords; i += 64) kset[w][i] = 0;
    if (lane == 0) hits[w] = counts[w] = 0;
    __syncthreads();
    const bool live = p < a.n_pairs;
    const uint64_t p1 = a.one_query ? 0 : p;
    const uint8_t *s1 = live ? a.seq1 + a.off1[p1] : nullptr;
    const uint8_t *s2 = live ? a.seq2 + a.off2[p] : nullptr;
    const int m = live ? int(a.off1[p1 + 1] - a.off1[p1]) : 0;
    const int n = live ? int(a.off2[p + 1] - a.off2[p]) : 0;
    const int k = a.k;
    // 3-mer histograms and the minimizer set of s1
    for (int i = lane; i + 2 < m; i += 64) {
        const uint32_t h = mix32((uint32_t(s1[i]) << 16) | (uint32_t(s1[i + 1]) << 8) | s1[i + 2]) >> 26;
        atomicAdd(&comp[w][0][h], 1u);
    }
    for (int i = lane; i + 2 < n; i += 64) {
        const uint32_t h = mix32((uint32_t(s2[i]) << 16) | (uint32_t(s2[i + 1]) << 8) | s2[i + 2]) >> 26;
        atomicAdd(&comp[w][1][h], 1u);
    }
    for_each_minimizer(s1, m, a, [&](uint32_t h) {
        h &= kSetWords * 32 - 1;
        atomicOr(&kset[w][h >> 5], 1u << (h & 31));
    });
    __syncthreads();
    uint32_t local_hits = 0, local_count = 0;
    for_each_minimizer(s2, n, a, [&](uint32_t h) {
        h &= kSetWords * 32 - 1;
        local_hits += (kset[w][h >> 5] >> (h & 31)) & 1u;
        local_count += 1;
    });
    atomicAdd(&hits[w], local_hits);
    atomicAdd(&counts[w], local_count);
    __syncthreads();
    if (!live) return;
    // L1 distance of the normalised histograms, reduced over the wave
    const float t1 = m > 2 ? 1.0f / float(m - 2) : 0.0f, t2 = n > 2 ? 1.0f / float(n - 2) : 0.0f;
    float d = fabsf(float(comp[w][0][lane]) * t1 - float(comp[w][1][lane]) * t2);
    for (int off = 32; off > 0; off >>= 1) d += __shfl_xor(d, off);
    if (lane == 0) {
        float *f = a.out + p * F;
        const float mn = float(m < n ? m : n), mx = float(m > n ? m : n);
        f[0] = log2f(float(m) + 1.0f) / 16.0f;
        f[1] = log2f(float(n) + 1.0f) / 16.0f;
        f[2] = mx > 0 ? mn / mx : 0.0f;
        f[3] = (m > 2 && n > 2) ? d * 0.5f : 1.0f;
        const float cont = (m >= k && n >= k && counts[w] > 0) ? float(hits[w]) / float(counts[w]) : 0.0f;
        f[4] = cont;
        f[5] = cont > 0 ? powf(cont, 1.0f / float(k)) : 0.0f;
        f[6] = a.smax / 16.0f;
        f[7] = 1.0f;
    }
}

// One thread carries one sample through the network; parameters staged in LDS.
__device__ __forceinline__ void load_params(float *sp, const float *params)
{
    for (int i = threadIdx.x; i < NPARAM; i += blockDim.x) sp[i] = params[i];
    __syncthreads();
}

__global__ __launch_bounds__(256) void forward_kernel(const float *params, const float *x, uint64_t n, float *y)
{
    __shared__ float sp[NPARAM];
    load_params(sp, params);
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float in[F], h1[HID], h2[HID];
#pragma unroll
    for (int f = 0; f < F; ++f) in[f] = x[i * F + f];
#pragma unroll
    for (int h = 0; h < HID; ++h) {
        float acc = sp[OFF_B1 + h];
#pragma unroll
        for (int f = 0; f < F; ++f) acc += in[f] * sp[OFF_W1 + f * HID + h];
        h1[h] = acc > 0.0f ? acc : 0.0f;
    }
#pragma unroll
    for (int h = 0; h < HID; ++h) {
        float acc = sp[OFF_B2 + h];
#pragma unroll 8
        for (int g = 0; g < HID; ++g) acc += h1[g] * sp[OFF_W2 + g * HID + h];
        h2[h] = acc > 0.0f ? acc : 0.0f;
    }
#pragma unroll
    for (int o = 0; o < OUT; ++o) {
        float acc = sp[OFF_B3 + o];
#pragma unroll
        for (int h = 0; h < HID; ++h) acc += h2[h] * sp[OFF_W3 + h * OUT + o];
        y[i * OUT + o] = acc;
    }
}

// Sum of v over the wave's 64 lanes, returned in every lane: running sums along each row of 16
// lanes by DPP row shifts (lane 15 of a row ends with its row's sum; lanes shifted in from outside
// the row add 0), then the four rows' sums.  Every lane of the wave must be active.
__device__ __forceinline__ float wave_sum(float v)
{
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xF, 0xF, false));   // row_shr:1
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x112, 0xF, 0xF, false));   // row_shr:2
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x114, 0xF, 0xF, false));   // row_shr:4
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x118, 0xF, 0xF, false));   // row_shr:8
    const int b = __float_as_int(v);
    return (__int_as_float(__builtin_amdgcn_readlane(b, 15)) + __int_as_float(__builtin_amdgcn_readlane(b, 31))) +
           (__int_as_float(__builtin_amdgcn_readlane(b, 47)) + __int_as_float(__builtin_amdgcn_readlane(b, 63)));
}

// Mean squared error (weighted per output) and its gradient.  Each wave sums every gradient entry
// over its 64 samples in registers (wave_sum), one lane adds the sum to the block's LDS copy, and
// the block adds that to the global gradient once.  Threads past the batch carry zero weight, so
// they add nothing but keep every wave whole for the lane sums.
__global__ __launch_bounds__(256) void loss_grad_kernel(const float *params, const float *x, const float *target,
                                                        const float *weight, uint64_t n, float inv_n, float *grad,
                                                        float *loss)
{
    __shared__ float sp[NPARAM];
    __shared__ float sg[NPARAM];
    __shared__ float sloss;
    for (int i = threadIdx.x; i < NPARAM; i += blockDim.x) sg[i] = 0.0f;
    if (threadIdx.x == 0) sloss = 0.0f;
    load_params(sp, params);   // ends with a barrier: sg and sloss are cleared too
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool live = i < n;
    const bool lead = threadIdx.x % 64 == 0;
    auto add = [&](float *dst, float v) {
        v = wave_sum(v);
        if (lead) atomicAdd(dst, v);
    };
    float in[F], h1[HID], h2[HID], dy[OUT], dh2[HID];
#pragma unroll
    for (int f = 0; f < F; ++f) in[f] = live ? x[i * F + f] : 0.0f;
#pragma unroll
    for (int h = 0; h < HID; ++h) {
        float acc = sp[OFF_B1 + h];
#pragma unroll
        for (int f = 0; f < F; ++f) acc += in[f] * sp[OFF_W1 + f * HID + h];
        h1[h] = acc > 0.0f ? acc : 0.0f;
    }
#pragma unroll
    for (int h = 0; h < HID; ++h) {
        float acc = sp[OFF_B2 + h];
#pragma unroll 8
        for (int g = 0; g < HID; ++g) acc += h1[g] * sp[OFF_W2 + g * HID + h];
        h2[h] = acc > 0.0f ? acc : 0.0f;
    }
    float l = 0.0f;
#pragma unroll
    for (int o = 0; o < OUT; ++o) {
        float acc = sp[OFF_B3 + o];
#pragma unroll
        for (int h = 0; h < HID; ++h) acc += h2[h] * sp[OFF_W3 + h * OUT + o];
        const float wt = live ? (weight ? weight[i * OUT + o] : 1.0f) : 0.0f;
        const float e = live ? acc - target[i * OUT + o] : 0.0f;
        l += wt * e * e;
        dy[o] = 2.0f * wt * e * inv_n;
    }
    add(&sloss, l * inv_n);
    // layer 3
#pragma unroll
    for (int o = 0; o < OUT; ++o) add(&sg[OFF_B3 + o], dy[o]);
#pragma unroll
    for (int h = 0; h < HID; ++h) {
        float g = 0.0f;
#pragma unroll
        for (int o = 0; o < OUT; ++o) {
            add(&sg[OFF_W3 + h * OUT + o], h2[h] * dy[o]);
            g += sp[OFF_W3 + h * OUT + o] * dy[o];
        }
        dh2[h] = h2[h] > 0.0f ? g : 0.0f;
    }
    // layer 2
#pragma unroll
    for (int h = 0; h < HID; ++h) add(&sg[OFF_B2 + h], dh2[h]);
    float dh1[HID];
#pragma unroll
    for (int g = 0; g < HID; ++g) {
        float acc = 0.0f;
#pragma unroll 8
        for (int h = 0; h < HID; ++h) {
            add(&sg[OFF_W2 + g * HID + h], h1[g] * dh2[h]);
            acc += sp[OFF_W2 + g * HID + h] * dh2[h];
        }
        dh1[g] = h1[g] > 0.0f ? acc : 0.0f;
    }
    // layer 1
#pragma unroll
    for (int h = 0; h < HID; ++h) {
        add(&sg[OFF_B1 + h], dh1[h]);
#pragma unroll
        for (int f = 0; f < F; ++f) add(&sg[OFF_W1 + f * HID + h], in[f] * dh1[h]);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < NPARAM; j += blockDim.x)
        if (sg[j] != 0.0f) atomicAdd(&grad[j], sg[j]);
    if (threadIdx.x == 0) atomicAdd(loss, sloss);
}

__global__ void fill_kernel(float *p, uint32_t n, float v)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void adam_kernel(float *params, const float *grad, float *m1, float *m2, uint32_t n, float lr, float b1,
                            float b2, float eps, float bc1, float bc2)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = grad[i];
    const float a = b1 * m1[i] + (1.0f - b1) * g;
    const float b = b2 * m2[i] + (1.0f - b2) * g * g;
    m1[i] = a;
    m2[i] = b;
    params[i] -= lr * (a / bc1) / (sqrtf(b / bc2) + eps);
}

hipError_t launch_features(const FeatureArgs &a, hipStream_t st)
{
    if (a.n_pairs == 0) return hipSuccess;
    hipLaunchKernelGGL(features_kernel, dim3(uint32_t((a.n_pairs + 3) / 4)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_forward(const float *params, const float *x, uint64_t n, float *y, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(forward_kernel, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, st, params, x, n, y);
    return hipGetLastError();
}

hipError_t launch_loss_grad(const float *params, const float *x, const float *target, const float *weight, uint64_t n,
                            float *grad, float *loss, hipStream_t st)
{
    hipLaunchKernelGGL(fill_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, st, grad, uint32_t(NPARAM), 0.0f);
    hipLaunchKernelGGL(fill_kernel, dim3(1), dim3(32), 0, st, loss, 1u, 0.0f);
    if (n == 0) return hipGetLastError();
    hipLaunchKernelGGL(loss_grad_kernel, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, st, params, x, target, weight,
                       n, 1.0f / float(n), grad, loss);
    return hipGetLastError();
}

hipError_t launch_adam(float *params, const float *grad, float *m1, float *m2, float lr, int step, hipStream_t st)
{
    const float b1 = 0.9f, b2 = 0.999f;
    float bc1 = 1.0f, bc2 = 1.0f;
    for (int i = 0; i < step; ++i) { bc1 *= b1; bc2 *= b2; }
    hipLaunchKernelGGL(adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, st, params, grad, m1, m2,
                       uint32_t(NPARAM), lr, b1, b2, 1e-8f, 1.0f - bc1, 1.0f - bc2);
    return hipGetLastError();
}

} // namespace pa
