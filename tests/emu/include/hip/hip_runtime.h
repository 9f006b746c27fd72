// Kernel side of the CPU wave emulator (test infrastructure only).
//
// Every lane of a workgroup runs the kernel body as its own fiber.  Lanes of a
// wave meet at every cross-lane operation (ballot, shuffles, readlane,
// readfirstlane, the wave barrier), which the emulator resolves once all live
// lanes of the wave have arrived at the same operation from the same source
// line; a wave whose lanes arrive at different operations is reported as
// divergent.  __syncthreads releases once every wave that has not finished
// waits at it.  Waves run one after another between barriers, or
// (RPS_EMU_SCHEDULE=interleave|random) one cross-lane operation at a time in
// turn, in a fixed or a seeded random order.  Workgroups of a launch run one after another, so __shared__
// variables become function statics shared by the lanes of the running
// workgroup.  A stuck launch (a lane looping without reaching an operation,
// or a barrier no other wave reaches) is reported with the position of every
// lane and ends the process.
#ifndef RPS_EMU_HIP_RUNTIME_H
#define RPS_EMU_HIP_RUNTIME_H

#include <cstdint>
#include <cstring>
#include <type_traits>
#include <utility>

#include "hip_runtime_api.h"

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline __attribute__((always_inline))
#define __launch_bounds__(...)
// All __shared__ variables share one section, which the emulator can fill
// with garbage before each workgroup (RPS_EMU_POISON_LDS=1) so that a read
// of LDS no lane of the workgroup wrote shows up.
#define __shared__ static __attribute__((section("rps_emu_lds")))
#define __align__(n) __attribute__((aligned(n)))

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned x_ = 1, unsigned y_ = 1, unsigned z_ = 1) : x(x_), y(y_), z(z_) {}
};

namespace emu {

enum Op : int {
    kRunnable = 0,
    kBallot,
    kShfl,
    kShflUp,
    kShflXor,
    kReadLane,
    kReadFirstLane,
    kRowBcast,
    kWaveSync,
    kBarrier,
    kExited,
};

struct Idx {
    unsigned x, y, z;
};

const Idx &thread_idx();
const Idx &block_idx();
const Idx &block_dim();
const Idx &grid_dim();
int lane_id();
// Lane n of row_newbcast:n (dpp_ctrl 0x150 + n).
int row_bcast_lane(int dpp_ctrl);
// A cross-lane operation: blocks the lane until its wave resolves it.
uint64_t wave_op(Op op, uint64_t value, int arg, int line);
void barrier(int line);

using Entry = void (*)(void *closure);
void launch(const char *name, dim3 grid, dim3 block, Entry entry, void *closure);

template <class F>
void invoke(void *closure) {
    (*static_cast<F *>(closure))();
}

template <class K, class... A>
void launch_kernel(const char *name, K kernel, dim3 grid, dim3 block, A... args) {
    auto body = [&]() { kernel(args...); };
    launch(name, grid, block, &invoke<decltype(body)>, &body);
}

template <class T>
inline uint64_t bits_of(T v) {
    static_assert(sizeof(T) <= 8, "cross-lane values of at most 64 bits");
    uint64_t u = 0;
    std::memcpy(&u, &v, sizeof(T));
    return u;
}

template <class T>
inline T from_bits(uint64_t u) {
    T v;
    std::memcpy(&v, &u, sizeof(T));
    return v;
}

}  // namespace emu

#define threadIdx (::emu::thread_idx())
#define blockIdx (::emu::block_idx())
#define blockDim (::emu::block_dim())
#define gridDim (::emu::grid_dim())

#define __syncthreads() ::emu::barrier(__LINE__)

#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...) \
    ::emu::launch_kernel(#kernel, &kernel, dim3(grid), dim3(block), __VA_ARGS__)

// One workgroup at a time runs on the emulator.
#define hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, ...) (*(blocks) = 1, hipSuccess)

inline uint32_t __umulhi(uint32_t a, uint32_t b) {
    return static_cast<uint32_t>((static_cast<uint64_t>(a) * b) >> 32);
}
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __popc(unsigned x) { return __builtin_popcount(x); }
inline int __ffsll(long long x) { return __builtin_ffsll(x); }

#define __ballot(pred) \
    (static_cast<unsigned long long>(::emu::wave_op(::emu::kBallot, (pred) ? 1u : 0u, 0, __LINE__)))
#define __shfl(v, src, ...) \
    (::emu::from_bits<std::decay_t<decltype(v)>>( \
        ::emu::wave_op(::emu::kShfl, ::emu::bits_of(v), (src), __LINE__)))
#define __shfl_up(v, d, ...) \
    (::emu::from_bits<std::decay_t<decltype(v)>>( \
        ::emu::wave_op(::emu::kShflUp, ::emu::bits_of(v), (d), __LINE__)))
#define __shfl_xor(v, d, ...) \
    (::emu::from_bits<std::decay_t<decltype(v)>>( \
        ::emu::wave_op(::emu::kShflXor, ::emu::bits_of(v), (d), __LINE__)))

#define __builtin_amdgcn_readfirstlane(v) \
    (static_cast<uint32_t>(::emu::wave_op(::emu::kReadFirstLane, static_cast<uint32_t>(v), 0, __LINE__)))
#define __builtin_amdgcn_readlane(v, l) \
    (static_cast<uint32_t>(::emu::wave_op(::emu::kReadLane, static_cast<uint32_t>(v), (l), __LINE__)))
// DPP row broadcast, the one DPP control the kernels use (row_newbcast:n,
// dpp_ctrl 0x150 + n): every lane takes lane n of its 16-lane row.  Any
// other control ends the process.
#define __builtin_amdgcn_mov_dpp(v, ctrl, row_mask, bank_mask, bound_ctrl) \
    (static_cast<int>(::emu::wave_op(::emu::kRowBcast, static_cast<uint32_t>(v), \
                                     ::emu::row_bcast_lane(ctrl), __LINE__)))
#define __builtin_amdgcn_wave_barrier() ((void)::emu::wave_op(::emu::kWaveSync, 0, 0, __LINE__))
#define __builtin_amdgcn_fence(order, scope) ((void)0)
#define __builtin_amdgcn_s_sleep(n) ((void)0)
#define __builtin_amdgcn_s_setprio(n) ((void)0)

inline uint32_t __builtin_amdgcn_mbcnt_lo(uint32_t mask, uint32_t acc) {
    const int l = ::emu::lane_id();
    const uint32_t below = l >= 32 ? mask : (l ? mask & ((1u << l) - 1u) : 0u);
    return acc + static_cast<uint32_t>(__builtin_popcount(below));
}
inline uint32_t __builtin_amdgcn_mbcnt_hi(uint32_t mask, uint32_t acc) {
    const int l = ::emu::lane_id();
    const uint32_t below = l < 32 ? 0u : (l > 32 ? mask & ((1u << (l - 32)) - 1u) : 0u);
    return acc + static_cast<uint32_t>(__builtin_popcount(below));
}
inline uint32_t __builtin_amdgcn_alignbit(uint32_t hi, uint32_t lo, uint32_t shift) {
    return static_cast<uint32_t>((((static_cast<uint64_t>(hi) << 32) | lo) >> (shift & 31)));
}
// Three-input bitwise function given by its 8-entry truth table: bit
// (a << 2 | b << 1 | c) of imm is the result for input bits a, b, c.
inline uint32_t __builtin_amdgcn_bitop3_b32(uint32_t a, uint32_t b, uint32_t c, unsigned imm) {
    uint32_t r = 0;
    for (int k = 0; k < 8; ++k) {
        if (imm & (1u << k)) {
            const uint32_t ma = (k & 4) ? a : ~a;
            const uint32_t mb = (k & 2) ? b : ~b;
            const uint32_t mc = (k & 1) ? c : ~c;
            r |= ma & mb & mc;
        }
    }
    return r;
}

// Atomics: one host thread runs every lane, so plain read-modify-writes.
template <class T>
inline T atomicAdd(T *p, T v) {
    T old = *p;
    *p = old + v;
    return old;
}
template <class T>
inline T atomicMin(T *p, T v) {
    T old = *p;
    *p = v < old ? v : old;
    return old;
}
template <class T>
inline T atomicMax(T *p, T v) {
    T old = *p;
    *p = v > old ? v : old;
    return old;
}
template <class T>
inline T atomicOr(T *p, T v) {
    T old = *p;
    *p = old | v;
    return old;
}
template <class T>
inline T atomicExch(T *p, T v) {
    T old = *p;
    *p = v;
    return old;
}
#define __HIP_MEMORY_SCOPE_AGENT 0
#define __HIP_MEMORY_SCOPE_WORKGROUP 1
#define __hip_atomic_load(p, order, scope) (*(p))
#define __hip_atomic_store(p, v, order, scope) ((void)(*(p) = (v)))

#endif
