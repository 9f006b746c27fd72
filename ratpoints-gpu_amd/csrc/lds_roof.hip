// LDS read-bandwidth probe for the roofline of the sieve kernel.
//
// Every wave reads 64 consecutive 8-byte words (lane l reads word s + l, s
// wave-uniform) -- the sieve kernel's access pattern -- and XORs them into a
// register; the loop offsets vary every iteration so nothing is hoisted.
// Prints the achieved LDS read bandwidth for: 8-byte reads per lane; 16-byte
// reads per lane at an 8-byte-aligned base and at a 16-byte-aligned base;
// and pairs of 8-byte reads 64 words apart (the dense stage's two words per
// lane, one ds_read2st64_b64).
// `lds_roof copy` instead runs one plain 256 MiB device-to-device copy kernel,
// the reference for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters.
//
// Usage: lds_roof [iterations] | lds_roof copy
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <string>

namespace {

constexpr int kWords = 8192;  // 64 KiB of LDS
constexpr int kThreads = 512;

// VEC words per lane, contiguous (STRIDE 1) or 64 words apart (STRIDE 64);
// ALIGN 2 rounds the wave's base down to an even word (16-byte aligned).
template <int VEC, int STRIDE = 1, int ALIGN = 1>
__global__ __launch_bounds__(kThreads) void lds_read_kernel(
    int iterations, unsigned long long *out) {
    constexpr int kPad = 64 * VEC * STRIDE;
    __shared__ __attribute__((aligned(16))) uint64_t buf[kWords + kPad];
    for (int i = threadIdx.x; i < kWords + kPad; i += kThreads) {
        buf[i] = 0x9E3779B97F4A7C15ULL * (i + 1);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    uint32_t s = __builtin_amdgcn_readfirstlane((threadIdx.x >> 6) * 977u);
    // Each loaded 64-bit word is folded into one 32-bit accumulator with a
    // single three-input XOR (v_bitop3_b32), so the VALU work per byte read
    // stays well below the LDS rate being measured.
    uint32_t acc[VEC];
    for (int v = 0; v < VEC; ++v) {
        acc[v] = 0;
    }
    for (int it = 0; it < iterations; ++it) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t base = ((s + 131u * k) & (kWords - 1)) & ~(ALIGN - 1u);
            const int step = STRIDE == 1 ? VEC : 1;
            const uint64_t *p = buf + base + lane * step;
            if (ALIGN == 2) {
                p = static_cast<const uint64_t *>(__builtin_assume_aligned(p, 16));
            }
#pragma unroll
            for (int v = 0; v < VEC; ++v) {
                const uint64_t w = p[v * STRIDE];
                acc[v] = __builtin_amdgcn_bitop3_b32(acc[v], static_cast<uint32_t>(w),
                                                     static_cast<uint32_t>(w >> 32), 0x96);
            }
        }
        s += 1031u;
    }
    uint64_t r = 0;
    for (int v = 0; v < VEC; ++v) {
        r ^= acc[v];
    }
    if (r == 0x1234567ULL) {
        out[blockIdx.x] = r;  // never true in practice; keeps the loads live
    }
}

template <int VEC, int STRIDE = 1, int ALIGN = 1>
double measure(int iterations, int blocks, unsigned long long *d_out) {
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    hipLaunchKernelGGL((lds_read_kernel<VEC, STRIDE, ALIGN>), dim3(blocks), dim3(kThreads), 0, 0,
                       10, d_out);
    (void)hipDeviceSynchronize();
    (void)hipEventRecord(a);
    hipLaunchKernelGGL((lds_read_kernel<VEC, STRIDE, ALIGN>), dim3(blocks), dim3(kThreads), 0, 0,
                       iterations, d_out);
    (void)hipEventRecord(b);
    (void)hipEventSynchronize(b);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    const double bytes = 8.0 * VEC * 64 * 8 * (kThreads / 64) * double(blocks) * iterations;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return bytes / (ms * 1e-3) / 1e9;
}

__global__ void copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                            size_t n) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n;
         i += size_t(gridDim.x) * blockDim.x) {
        dst[i] = src[i];
    }
}

int run_copy() {
    const size_t bytes = size_t(256) << 20;
    const size_t n = bytes / sizeof(uint4);
    uint4 *src = nullptr, *dst = nullptr;
    if (hipMalloc(&src, bytes) != hipSuccess || hipMalloc(&dst, bytes) != hipSuccess) {
        return 1;
    }
    (void)hipMemset(src, 1, bytes);
    (void)hipDeviceSynchronize();
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, 0, src, dst, n);
    (void)hipDeviceSynchronize();
    std::printf("copy_kernel bytes_read=%zu bytes_written=%zu (KiB %zu each)\n", bytes, bytes,
                bytes / 1024);
    (void)hipFree(src);
    (void)hipFree(dst);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc > 1 && std::string(argv[1]) == "copy") {
        return run_copy();
    }
    const int iterations = argc > 1 ? std::atoi(argv[1]) : 20000;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
        std::fprintf(stderr, "no HIP device\n");
        return 1;
    }
    const int blocks = prop.multiProcessorCount * 2;
    unsigned long long *d_out = nullptr;
    if (hipMalloc(&d_out, sizeof(unsigned long long) * blocks) != hipSuccess) {
        return 1;
    }
    const double b64 = measure<1>(iterations, blocks, d_out);
    const double b128 = measure<2>(iterations, blocks, d_out);
    const double b128a = measure<2, 1, 2>(iterations, blocks, d_out);
    const double st64 = measure<2, 64>(iterations, blocks, d_out);
    const double per_clk = 1e9 / (prop.multiProcessorCount * prop.clockRate * 1e3);
    std::printf("device=%s cus=%d clock_khz=%d\n", prop.name, prop.multiProcessorCount,
                prop.clockRate);
    std::printf("lds_read_b64_GBs=%.1f lds_read_b128_GBs=%.1f lds_read_b128_aligned_GBs=%.1f "
                "lds_read2st64_b64_GBs=%.1f\n", b64, b128, b128a, st64);
    std::printf("bytes_per_clk_per_cu_b64=%.1f bytes_per_clk_per_cu_b128=%.1f "
                "bytes_per_clk_per_cu_b128_aligned=%.1f bytes_per_clk_per_cu_read2st64=%.1f\n",
                b64 * per_clk, b128 * per_clk, b128a * per_clk, st64 * per_clk);
    (void)hipFree(d_out);
    return 0;
}
