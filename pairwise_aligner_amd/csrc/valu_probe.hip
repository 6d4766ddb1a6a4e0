// valu_probe.hip -- issue-rate probe for the instructions the DP kernels are built from
// (v_pk_add_i16 and v_pk_max_i16 with and without dependencies, v_pk_max_u16, v_add_u32,
// v_max_i32, v_xor_b32, v_perm_b32, row_shr DPP moves) at
// 1, 2, 4 and 8 waves per SIMD.  Each wave runs `iters` x 16 instructions between two reads of
// the shader clock (s_memtime) and the 100 MHz real-time clock (s_memrealtime); the probe prints
// one JSON line per configuration with
//   cycles_per_instr  shader cycles per instruction of one wave,
//   sclk_mhz          shader clock during the run (cycles over real time),
//   lane_ops_t        lane-instructions per second over the whole GPU (HIP-event kernel time),
//                     packed ops counted once (each does two 16-bit operations).
// Built by pairwise_aligner_amd/build.py; run by `python bench.py --valu-probe`.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <vector>

#define PROBE_CHECK(x)                                                                           \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) {                                                                  \
            std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);  \
            return 1;                                                                            \
        }                                                                                        \
    } while (0)

enum Mode { kPkAdd, kPkAddDep, kPkAddDep2, kPkAddDep4, kPkMax, kAddU32, kDppShr, kMaxI32, kXorB32, kPermB32, kPkMaxU16 };

template <int MODE>
__global__ __launch_bounds__(64) void probe_kernel(uint32_t *out, uint64_t *clk, int iters, uint32_t seed)
{
    uint32_t a[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) a[i] = seed + threadIdx.x * 16 + i;
    const uint32_t b = seed ^ 0x00010001u;
    const uint64_t t0 = __builtin_amdgcn_s_memtime();
    const uint64_t r0 = __builtin_amdgcn_s_memrealtime();
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            // register operands only: the asm blocks are plain VALU instructions, no memory access
            if constexpr (MODE == kPkAdd) asm volatile("v_pk_add_i16 %0, %0, %1 clamp" : "+v"(a[i]) : "v"(b));
            if constexpr (MODE == kPkAddDep) asm volatile("v_pk_add_i16 %0, %0, %1 clamp" : "+v"(a[0]) : "v"(b));
            if constexpr (MODE == kPkAddDep2) asm volatile("v_pk_add_i16 %0, %0, %1 clamp" : "+v"(a[i & 1]) : "v"(b));
            if constexpr (MODE == kPkAddDep4) asm volatile("v_pk_add_i16 %0, %0, %1 clamp" : "+v"(a[i & 3]) : "v"(b));
            if constexpr (MODE == kPkMax) asm volatile("v_pk_max_i16 %0, %0, %1" : "+v"(a[i]) : "v"(b));
            if constexpr (MODE == kAddU32) asm volatile("v_add_u32 %0, %0, %1" : "+v"(a[i]) : "v"(b));
            if constexpr (MODE == kMaxI32) asm volatile("v_max_i32 %0, %0, %1" : "+v"(a[i]) : "v"(b));
            if constexpr (MODE == kXorB32) asm volatile("v_xor_b32 %0, %0, %1" : "+v"(a[i]) : "v"(b));
            if constexpr (MODE == kPermB32) asm volatile("v_perm_b32 %0, %0, %1, %1" : "+v"(a[i]) : "v"(b));
            if constexpr (MODE == kPkMaxU16) asm volatile("v_pk_max_u16 %0, %0, %1" : "+v"(a[i]) : "v"(b));
            if constexpr (MODE == kDppShr)
                a[i] = uint32_t(__builtin_amdgcn_update_dpp(int(a[i]), int(a[(i + 1) & 15]), 0x111, 0xF, 0xF, false));
        }
    }
    const uint64_t t1 = __builtin_amdgcn_s_memtime();
    const uint64_t r1 = __builtin_amdgcn_s_memrealtime();
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) s ^= a[i];
    out[blockIdx.x * 64 + threadIdx.x] = s;
    if (threadIdx.x == 0) {
        clk[2 * blockIdx.x] = t1 - t0;
        clk[2 * blockIdx.x + 1] = r1 - r0;
    }
}

constexpr int kSimds = 1024;        // 256 CUs x 4 SIMDs
constexpr int kMaxWavesPerSimd = 8;

template <int MODE>
int run(const char *name, int waves_per_simd, int iters, uint32_t *out, uint64_t *clk)
{
    const int blocks = kSimds * waves_per_simd;   // one wave per block
    hipEvent_t e0, e1;
    PROBE_CHECK(hipEventCreate(&e0));
    PROBE_CHECK(hipEventCreate(&e1));
    probe_kernel<MODE><<<blocks, 64>>>(out, clk, iters, 1);   // warm-up
    PROBE_CHECK(hipGetLastError());
    PROBE_CHECK(hipDeviceSynchronize());
    PROBE_CHECK(hipEventRecord(e0));
    probe_kernel<MODE><<<blocks, 64>>>(out, clk, iters, 2);
    PROBE_CHECK(hipGetLastError());
    PROBE_CHECK(hipEventRecord(e1));
    PROBE_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    PROBE_CHECK(hipEventElapsedTime(&ms, e0, e1));
    std::vector<uint64_t> h(2 * size_t(blocks));
    PROBE_CHECK(hipMemcpy(h.data(), clk, h.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
    double cyc = 0, rt = 0;
    for (int i = 0; i < blocks; ++i) { cyc += double(h[2 * i]); rt += double(h[2 * i + 1]); }
    cyc /= blocks;
    rt /= blocks;
    const double instr = double(iters) * 16;
    std::printf("{\"op\": \"%s\", \"waves_per_simd\": %d, \"cycles_per_instr\": %.3f, \"sclk_mhz\": %.0f, "
                "\"kernel_ms\": %.4f, \"lane_ops_t\": %.3f}\n",
                name, waves_per_simd, cyc / instr, cyc / (rt / 100.0), ms,
                instr * blocks * 64 / (double(ms) * 1e-3) / 1e12);
    PROBE_CHECK(hipEventDestroy(e0));
    PROBE_CHECK(hipEventDestroy(e1));
    return 0;
}

int main()
{
    uint32_t *out = nullptr;
    uint64_t *clk = nullptr;
    const size_t max_blocks = size_t(kSimds) * kMaxWavesPerSimd;
    PROBE_CHECK(hipMalloc(&out, max_blocks * 64 * sizeof(uint32_t)));
    PROBE_CHECK(hipMalloc(&clk, max_blocks * 2 * sizeof(uint64_t)));
    const int iters = 4096;
    for (int w = 1; w <= kMaxWavesPerSimd; w *= 2) {
        if (run<kPkAdd>("v_pk_add_i16", w, iters, out, clk)) return 1;
        if (run<kPkAddDep>("v_pk_add_i16 one chain", w, iters, out, clk)) return 1;
        if (run<kPkAddDep2>("v_pk_add_i16 two chains", w, iters, out, clk)) return 1;
        if (run<kPkAddDep4>("v_pk_add_i16 four chains", w, iters, out, clk)) return 1;
        if (run<kPkMax>("v_pk_max_i16", w, iters, out, clk)) return 1;
        if (run<kAddU32>("v_add_u32", w, iters, out, clk)) return 1;
        if (run<kDppShr>("v_mov_b32_dpp row_shr:1", w, iters, out, clk)) return 1;
        if (run<kMaxI32>("v_max_i32", w, iters, out, clk)) return 1;
        if (run<kXorB32>("v_xor_b32", w, iters, out, clk)) return 1;
        if (run<kPermB32>("v_perm_b32", w, iters, out, clk)) return 1;
        if (run<kPkMaxU16>("v_pk_max_u16", w, iters, out, clk)) return 1;
    }
    PROBE_CHECK(hipFree(out));
    PROBE_CHECK(hipFree(clk));
    return 0;
}
