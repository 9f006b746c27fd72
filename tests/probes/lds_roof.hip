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
// `lds_roof rows` times the wide-row dense stage's access pattern: one
// workgroup of 16 waves per CU, every wave reading per iteration NL rows of
// 16-byte elements from LDS (ds_read_b128 at (off_r + lane) * 16) and NG rows
// from a per-workgroup global buffer that stays in the vector L1 cache
// (global_load_dwordx4), row offsets advancing by 128 mod p_r per iteration:
// three scalar instructions per row, or (offsets=vgpr) one VGPR holding
// every row's offset, advanced by three vector instructions and read back
// with v_readlane.
//
// Usage: lds_roof [iterations] | lds_roof copy | lds_roof rows [iterations]
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

constexpr int kRowSlot = 576;      // elements per row slot (p + 64 <= 573)
constexpr int kRowThreads = 1024;  // 16 waves
constexpr int kMaxRows = 14;
__constant__ uint32_t c_row_p[kMaxRows] = {509, 503, 499, 491, 487, 479, 467,
                                           463, 461, 457, 449, 443, 439, 433};

template <int NL, int NG, bool VOFF = false>
__global__ __launch_bounds__(kRowThreads) void rows_kernel(int iterations, const uint4 *grows,
                                                           unsigned *out) {
    static_assert(!VOFF || NG == 0, "vector offsets are probed on LDS rows only");
    __shared__ __attribute__((aligned(16))) uint4 rows[(NL > 0 ? NL : 1) * kRowSlot];
    for (int i = threadIdx.x; i < NL * kRowSlot; i += kRowThreads) {
        rows[i] = make_uint4(0x9E3779B9u * (i + 1), 0x85EBCA6Bu * (i + 3),
                             0xC2B2AE35u * (i + 5), 0x27D4EB2Fu * (i + 7));
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const uint4 *g = grows + static_cast<size_t>(blockIdx.x) * (NG > 0 ? NG : 1) * kRowSlot;
    uint32_t off[NL + NG];
    uint32_t p[NL + NG];
#pragma unroll
    for (int r = 0; r < NL + NG; ++r) {
        p[r] = __builtin_amdgcn_readfirstlane(c_row_p[r]);
        off[r] = __builtin_amdgcn_readfirstlane((wave * 131u + r * 37u) % p[r]);
    }
    // Offsets=vgpr: lane r < NL holds row r's offset and prime.
    const uint32_t lp = lane < NL ? c_row_p[lane] : 128u;
    uint32_t voff = lane < NL ? (wave * 131u + lane * 37u) % lp : 0u;
    uint32_t a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    for (int it = 0; it < iterations; ++it) {
        uint4 v[NL + NG];
        if (VOFF) {
#pragma unroll
            for (int r = 0; r < NL; ++r) {
                off[r] = __builtin_amdgcn_readlane(voff, r);
            }
        }
#pragma unroll
        for (int r = 0; r < NL; ++r) {
            v[r] = rows[r * kRowSlot + off[r] + lane];
        }
#pragma unroll
        for (int r = 0; r < NG; ++r) {
            v[NL + r] = g[r * kRowSlot + off[NL + r] + lane];
        }
        if (VOFF) {
            const uint32_t next = voff + 128u;
            voff = next >= lp ? next - lp : next;
        } else {
#pragma unroll
            for (int r = 0; r < NL + NG; ++r) {
                const uint32_t next = off[r] + 128u;
                off[r] = next >= p[r] ? next - p[r] : next;
                asm volatile("" ::"s"(off[r]));
            }
        }
#pragma unroll
        for (int r = 0; r < NL + NG; ++r) {
            a0 = __builtin_amdgcn_bitop3_b32(a0, v[r].x, v[r].y, 0x96);
            a1 = __builtin_amdgcn_bitop3_b32(a1, v[r].z, v[r].w, 0x96);
        }
        a2 += a0;
        a3 ^= a1;
    }
    if ((a0 ^ a1 ^ a2 ^ a3) == 0x1234567u) {
        out[blockIdx.x] = a0;  // keeps the loads live; never true in practice
    }
}

template <int NL, int NG, bool VOFF = false>
void measure_rows(int iterations, int blocks, const uint4 *grows, unsigned *out,
                  double clk_hz) {
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    hipLaunchKernelGGL((rows_kernel<NL, NG, VOFF>), dim3(blocks), dim3(kRowThreads), 0, 0, 10,
                       grows, out);
    (void)hipDeviceSynchronize();
    (void)hipEventRecord(a);
    hipLaunchKernelGGL((rows_kernel<NL, NG, VOFF>), dim3(blocks), dim3(kRowThreads), 0, 0,
                       iterations, grows, out);
    (void)hipEventRecord(b);
    (void)hipEventSynchronize(b);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    const double wave_iters_per_cu = double(kRowThreads / 64) * iterations;
    const double clk = ms * 1e-3 * clk_hz;
    std::printf("lds_rows=%d global_rows=%d offsets=%s ms=%.3f "
                "clk_per_wave_iteration_per_cu=%.1f row_bytes_per_clk_per_cu=%.1f\n",
                NL, NG, VOFF ? "vgpr" : "sgpr", ms, clk / wave_iters_per_cu,
                wave_iters_per_cu * (NL + NG) * 1024.0 / clk);
    std::fflush(stdout);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
}

int run_rows(int iterations) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
        return 1;
    }
    const int blocks = prop.multiProcessorCount;
    const double clk_hz = prop.clockRate * 1e3;
    const size_t n = static_cast<size_t>(blocks) * 3 * kRowSlot;
    uint4 *grows = nullptr;
    unsigned *out = nullptr;
    if (hipMalloc(&grows, n * sizeof(uint4)) != hipSuccess
        || hipMalloc(&out, blocks * sizeof(unsigned)) != hipSuccess) {
        return 1;
    }
    (void)hipMemset(grows, 0x5A, n * sizeof(uint4));
    (void)hipDeviceSynchronize();
    std::printf("cus=%d clock_khz=%d iterations=%d\n", blocks, prop.clockRate, iterations);
    measure_rows<14, 0>(iterations, blocks, grows, out, clk_hz);  // warm-up pass
    measure_rows<14, 0>(iterations, blocks, grows, out, clk_hz);
    measure_rows<12, 0>(iterations, blocks, grows, out, clk_hz);
    measure_rows<12, 2>(iterations, blocks, grows, out, clk_hz);
    measure_rows<11, 3>(iterations, blocks, grows, out, clk_hz);
    measure_rows<0, 2>(iterations, blocks, grows, out, clk_hz);
    measure_rows<0, 3>(iterations, blocks, grows, out, clk_hz);
    measure_rows<14, 0, true>(iterations, blocks, grows, out, clk_hz);
    measure_rows<12, 0, true>(iterations, blocks, grows, out, clk_hz);
    measure_rows<14, 0>(iterations, blocks, grows, out, clk_hz);
    measure_rows<14, 0, true>(iterations, blocks, grows, out, clk_hz);
    (void)hipFree(grows);
    (void)hipFree(out);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc > 1 && std::string(argv[1]) == "copy") {
        return run_copy();
    }
    if (argc > 1 && std::string(argv[1]) == "rows") {
        return run_rows(argc > 2 ? std::atoi(argv[2]) : 20000);
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
