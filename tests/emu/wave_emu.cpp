// CPU wave emulator (test infrastructure only): fiber scheduler for the
// lanes of a workgroup, resolution of cross-lane operations and barriers,
// a watchdog for stuck launches, and the host runtime API on host memory.
//
// Schedule: the waves of a workgroup run one after another between barriers.
// A wave's lanes run in lane order up to their next cross-lane operation;
// once every live lane waits at the same operation of the same source line,
// the operation is resolved and the wave continues, until all its lanes wait
// at a barrier or have returned.  When every wave of the workgroup waits at a
// barrier (or has returned), the barrier opens.  Workgroups of a launch run
// one after another, every operation of the stream completes before the call
// that enqueues it returns.
#include <hip/hip_runtime.h>

#include <sys/mman.h>
#include <sys/time.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

extern "C" void rps_emu_switch(void **save_sp, void *to_sp);

// AddressSanitizer builds (RPS_EMU_ASAN) are told about every stack switch,
// so that it checks the lanes' accesses on their own stacks.
#if defined(__has_feature)
#if __has_feature(address_sanitizer)
#define RPS_EMU_ASAN 1
#include <sanitizer/common_interface_defs.h>
#endif
#endif

// Saves the callee-saved registers and the stack pointer of the running
// context into *save_sp and resumes the context saved at to_sp (x86-64 SysV).
asm(R"(
    .text
    .p2align 4
    .globl rps_emu_switch
    .hidden rps_emu_switch
    .type rps_emu_switch,@function
rps_emu_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
    .size rps_emu_switch, .-rps_emu_switch
)");

namespace emu {
namespace {

constexpr size_t kStackBytes = 256 * 1024;
constexpr int kWaveSize = 64;
constexpr int kMaxLanes = 1024;

struct Lane {
    void *sp = nullptr;
    char *stack = nullptr;  // kStackBytes, lowest page a guard
    Idx tid{};
    int lane = 0;
    int wave = 0;
    int op = kExited;
    int line = 0;
    uint64_t value = 0;
    int arg = 0;
    uint64_t result = 0;
};

struct Launch {
    std::vector<Lane> lanes = std::vector<Lane>(kMaxLanes);
    int n_lanes = 0;
    int n_waves = 0;
    Lane *current = nullptr;
    void *sched_sp = nullptr;
    Idx block{}, bdim{}, gdim{};
    Entry entry = nullptr;
    void *closure = nullptr;
    const char *name = "";
    bool active = false;
};

Launch g;

#if RPS_EMU_ASAN
extern void *g_sched_fake;
extern const void *g_sched_bottom;
extern size_t g_sched_size;
#endif

const char *op_name(int op) {
    switch (op) {
    case kRunnable: return "running";
    case kBallot: return "ballot";
    case kShfl: return "shfl";
    case kShflUp: return "shfl_up";
    case kShflXor: return "shfl_xor";
    case kReadLane: return "readlane";
    case kReadFirstLane: return "readfirstlane";
    case kRowBcast: return "row_newbcast";
    case kWaveSync: return "wave_barrier";
    case kBarrier: return "__syncthreads";
    case kExited: return "returned";
    }
    return "?";
}

// Where every wave is: lanes grouped by (operation, line).
void describe(FILE *out) {
    std::fprintf(out, "emu: launch %s, workgroup (%u,%u,%u) of (%u,%u,%u), %d waves\n", g.name,
                 g.block.x, g.block.y, g.block.z, g.gdim.x, g.gdim.y, g.gdim.z, g.n_waves);
    for (int w = 0; w < g.n_waves; ++w) {
        std::string line;
        int run_op = -2, run_line = -1, run_from = 0;
        auto flush = [&](int upto) {
            if (run_op == -2) {
                return;
            }
            char buf[96];
            std::snprintf(buf, sizeof buf, " lanes %d-%d: %s@%d;", run_from, upto - 1,
                          op_name(run_op), run_line);
            line += buf;
        };
        for (int l = 0; l < kWaveSize; ++l) {
            const int i = w * kWaveSize + l;
            const int op = i < g.n_lanes ? g.lanes[i].op : kExited;
            const int ln = i < g.n_lanes ? g.lanes[i].line : 0;
            if (op != run_op || ln != run_line) {
                flush(l);
                run_op = op;
                run_line = ln;
                run_from = l;
            }
        }
        flush(kWaveSize);
        std::fprintf(out, "emu:   wave %d:%s\n", w, line.c_str());
    }
    std::fflush(out);
}

[[noreturn]] void fatal(const char *what) {
    std::fprintf(stderr, "emu: %s\n", what);
    describe(stderr);
    std::_Exit(86);
}

void watchdog(int) {
    static const char msg[] = "emu: watchdog: launch did not finish in time\n";
    (void)!write(2, msg, sizeof msg - 1);
    if (g.active) {
        if (g.current) {
            std::fprintf(stderr, "emu: lane %d of wave %d was running (last operation at line %d)\n",
                         g.current->lane, g.current->wave, g.current->line);
        }
        describe(stderr);
    }
    std::_Exit(87);
}

int watchdog_seconds() {
    const char *s = std::getenv("RPS_EMU_WATCHDOG_S");
    const int v = s ? std::atoi(s) : 900;
    return v > 0 ? v : 900;
}

void arm_watchdog(bool on) {
    static bool installed = false;
    if (!installed) {
        struct sigaction sa {};
        sa.sa_handler = watchdog;
        sigaction(SIGALRM, &sa, nullptr);
        installed = true;
    }
    itimerval t{};
    if (on) {
        t.it_value.tv_sec = watchdog_seconds();
    }
    setitimer(ITIMER_REAL, &t, nullptr);
}

void yield(Lane *l, bool exiting);

[[noreturn]] void lane_main() {
#if RPS_EMU_ASAN
    __sanitizer_finish_switch_fiber(nullptr, &g_sched_bottom, &g_sched_size);
#endif
    Lane *l = g.current;
    g.entry(g.closure);
    l->op = kExited;
    l->line = 0;
    yield(l, true);
    std::abort();  // an exited lane is never resumed
}

void reset_fiber(Lane &l) {
    if (!l.stack) {
        void *m = mmap(nullptr, kStackBytes, PROT_READ | PROT_WRITE,
                       MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
        if (m == MAP_FAILED) {
            std::fprintf(stderr, "emu: cannot map a fiber stack\n");
            std::_Exit(86);
        }
        mprotect(m, 4096, PROT_NONE);  // overflow guard
        l.stack = static_cast<char *>(m);
    }
    uintptr_t top = reinterpret_cast<uintptr_t>(l.stack + kStackBytes) & ~uintptr_t(15);
    auto *sp = reinterpret_cast<uintptr_t *>(top);
    *--sp = 0;  // return address slot of lane_main (never used)
    *--sp = reinterpret_cast<uintptr_t>(&lane_main);
    for (int k = 0; k < 6; ++k) {
        *--sp = 0;  // rbp, rbx, r12-r15
    }
    l.sp = sp;
}

#if RPS_EMU_ASAN
void *g_sched_fake = nullptr;
const void *g_sched_bottom = nullptr;
size_t g_sched_size = 0;
#endif

// Scheduler -> lane.
void run_lane(Lane &l) {
    g.current = &l;
#if RPS_EMU_ASAN
    __sanitizer_start_switch_fiber(&g_sched_fake, l.stack + 4096, kStackBytes - 4096);
#endif
    rps_emu_switch(&g.sched_sp, l.sp);
#if RPS_EMU_ASAN
    __sanitizer_finish_switch_fiber(g_sched_fake, nullptr, nullptr);
#endif
    g.current = nullptr;
}

// Lane -> scheduler; returns when the scheduler resumes the lane.
void yield(Lane *l, bool exiting) {
#if RPS_EMU_ASAN
    void *fake = nullptr;
    __sanitizer_start_switch_fiber(exiting ? nullptr : &fake, g_sched_bottom, g_sched_size);
#endif
    rps_emu_switch(&l->sp, g.sched_sp);
#if RPS_EMU_ASAN
    __sanitizer_finish_switch_fiber(fake, &g_sched_bottom, &g_sched_size);
#endif
}

void resolve(int w, int op) {
    Lane *L = &g.lanes[static_cast<size_t>(w) * kWaveSize];
    const int n = std::min(kWaveSize, g.n_lanes - w * kWaveSize);
    auto live = [&](int l) { return l < n && L[l].op != kExited; };
    uint64_t value[kWaveSize] = {};
    for (int l = 0; l < n; ++l) {
        if (live(l)) {
            value[l] = L[l].value;
        }
    }
    uint64_t ballot = 0;
    int first = -1;
    for (int l = 0; l < n; ++l) {
        if (!live(l)) {
            continue;
        }
        if (first < 0) {
            first = l;
        }
        if (value[l] & 1) {
            ballot |= 1ULL << l;
        }
    }
    for (int l = 0; l < n; ++l) {
        if (!live(l)) {
            continue;
        }
        uint64_t r = 0;
        switch (op) {
        case kBallot:
            r = ballot;
            break;
        case kShfl: {
            const int src = L[l].arg & (kWaveSize - 1);
            r = value[src];
            break;
        }
        case kShflUp: {
            const int src = l - L[l].arg;
            r = src >= 0 ? value[src] : value[l];
            break;
        }
        case kShflXor: {
            const int src = (l ^ L[l].arg) & (kWaveSize - 1);
            r = value[src];
            break;
        }
        case kReadLane:
            if (L[l].arg != L[first].arg || L[l].arg < 0 || L[l].arg >= kWaveSize) {
                fatal("readlane with a lane index that is not wave-uniform or out of range");
            }
            if (!live(L[l].arg)) {
                fatal("readlane of a lane that has returned");
            }
            r = value[L[l].arg];
            break;
        case kReadFirstLane:
            r = value[first];
            break;
        case kRowBcast: {
            if (L[l].arg != L[first].arg) {
                fatal("row_newbcast with a lane index that is not wave-uniform");
            }
            const int src = (l & ~15) + L[l].arg;
            if (!live(src)) {
                fatal("row_newbcast from a lane that is inactive or has returned");
            }
            r = value[src];
            break;
        }
        case kWaveSync:
            break;
        default:
            fatal("unknown cross-lane operation");
        }
        L[l].result = r;
        L[l].op = kRunnable;
    }
}

// Schedule and LDS poisoning, from the environment.
struct Options {
    int schedule = 0;  // 0 sequential, 1 interleave, 2 random
    bool poison = false;
    uint64_t seed = 1;
};

const Options &options() {
    static const Options o = [] {
        Options r;
        const char *s = std::getenv("RPS_EMU_SCHEDULE");
        if (s && std::string(s) == "interleave") {
            r.schedule = 1;
        } else if (s && std::string(s) == "random") {
            r.schedule = 2;
        }
        const char *p = std::getenv("RPS_EMU_POISON_LDS");
        r.poison = p && std::string(p) == "1";
        const char *seed = std::getenv("RPS_EMU_SEED");
        r.seed = seed ? std::strtoull(seed, nullptr, 10) : 1;
        return r;
    }();
    return o;
}

uint64_t g_rng = 0;

uint64_t next_random() {
    g_rng += 0x9E3779B97F4A7C15ULL;
    uint64_t z = g_rng;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

}  // namespace
}  // namespace emu

extern "C" char __start_rps_emu_lds[] __attribute__((weak));
extern "C" char __stop_rps_emu_lds[] __attribute__((weak));

namespace emu {
namespace {

void poison_lds() {
    if (!__start_rps_emu_lds || !__stop_rps_emu_lds) {
        return;
    }
    const int byte = static_cast<int>(next_random() & 0xFF);
    std::memset(__start_rps_emu_lds, byte, static_cast<size_t>(__stop_rps_emu_lds - __start_rps_emu_lds));
}

// Runs the runnable lanes of wave w to their next operation and settles
// the wave: 0 resolved an operation, 1 all lanes wait at a barrier, 2 all
// lanes have returned.
int step_wave(int w) {
    const int base = w * kWaveSize;
    const int n = std::min(kWaveSize, g.n_lanes - base);
    for (int l = 0; l < n; ++l) {
        if (g.lanes[base + l].op == kRunnable) {
            run_lane(g.lanes[base + l]);
        }
    }
    int live = 0, op = -1, line = 0;
    bool mixed = false;
    for (int l = 0; l < n; ++l) {
        const Lane &x = g.lanes[base + l];
        if (x.op == kExited) {
            continue;
        }
        if (live++ == 0) {
            op = x.op;
            line = x.line;
        } else if (x.op != op || x.line != line) {
            mixed = true;
        }
    }
    if (live == 0) {
        return 2;
    }
    if (mixed) {
        fatal("lanes of one wave wait at different operations (divergent "
              "cross-lane operation or barrier)");
    }
    if (op == kBarrier) {
        return 1;
    }
    resolve(w, op);
    return 0;
}

bool g_poison_launch = false;  // poison before each workgroup of this launch

void run_block() {
    if (g_poison_launch) {
        poison_lds();
    }
    std::vector<char> done(g.n_waves, 0), parked(g.n_waves, 0);
    for (int i = 0; i < g.n_lanes; ++i) {
        Lane &l = g.lanes[i];
        reset_fiber(l);
        const int flat = i;
        l.tid.x = static_cast<unsigned>(flat % g.bdim.x);
        l.tid.y = static_cast<unsigned>((flat / g.bdim.x) % g.bdim.y);
        l.tid.z = static_cast<unsigned>(flat / (g.bdim.x * g.bdim.y));
        l.lane = i % kWaveSize;
        l.wave = i / kWaveSize;
        l.op = kRunnable;
        l.line = 0;
    }
    std::vector<int> order(g.n_waves);
    for (int w = 0; w < g.n_waves; ++w) {
        order[w] = w;
    }
    for (;;) {
        if (options().schedule == 0) {
            // Each wave runs until it waits at a barrier or returns.
            for (int w = 0; w < g.n_waves; ++w) {
                while (!done[w] && !parked[w]) {
                    const int r = step_wave(w);
                    done[w] = r == 2;
                    parked[w] = r == 1;
                }
            }
        } else {
            // One operation per wave in turn (fixed or shuffled order),
            // until every wave waits at a barrier or has returned.
            for (bool moved = true; moved;) {
                moved = false;
                if (options().schedule == 2) {
                    for (int k = g.n_waves - 1; k > 0; --k) {
                        std::swap(order[k], order[next_random() % (k + 1)]);
                    }
                }
                for (int w : order) {
                    if (done[w] || parked[w]) {
                        continue;
                    }
                    const int r = step_wave(w);
                    done[w] = r == 2;
                    parked[w] = r == 1;
                    moved = true;
                }
            }
        }
        bool all_done = true;
        int bline = -1;
        for (int w = 0; w < g.n_waves; ++w) {
            if (done[w]) {
                continue;
            }
            all_done = false;
            for (int l = 0; l < std::min(kWaveSize, g.n_lanes - w * kWaveSize); ++l) {
                const Lane &x = g.lanes[w * kWaveSize + l];
                if (x.op == kExited) {
                    continue;
                }
                if (bline < 0) {
                    bline = x.line;
                } else if (x.line != bline) {
                    fatal("waves wait at different __syncthreads");
                }
            }
        }
        if (all_done) {
            return;
        }
        for (int w = 0; w < g.n_waves; ++w) {
            if (!parked[w]) {
                continue;
            }
            parked[w] = 0;
            for (int l = 0; l < std::min(kWaveSize, g.n_lanes - w * kWaveSize); ++l) {
                Lane &x = g.lanes[w * kWaveSize + l];
                if (x.op == kBarrier) {
                    x.op = kRunnable;
                }
            }
        }
    }
}

}  // namespace

const Idx &thread_idx() { return g.current->tid; }
const Idx &block_idx() { return g.block; }
const Idx &block_dim() { return g.bdim; }
const Idx &grid_dim() { return g.gdim; }
int lane_id() { return g.current->lane; }

int row_bcast_lane(int dpp_ctrl) {
    if (dpp_ctrl < 0x150 || dpp_ctrl > 0x15f) {
        fatal("DPP control other than row_newbcast");
    }
    return dpp_ctrl - 0x150;
}

uint64_t wave_op(Op op, uint64_t value, int arg, int line) {
    Lane *l = g.current;
    l->op = op;
    l->value = value;
    l->arg = arg;
    l->line = line;
    yield(l, false);
    return l->result;
}

void barrier(int line) {
    Lane *l = g.current;
    l->op = kBarrier;
    l->line = line;
    yield(l, false);
}

void launch(const char *name, dim3 grid, dim3 block, Entry entry, void *closure) {
    if (g.active) {
        std::fprintf(stderr, "emu: launch from inside a kernel\n");
        std::_Exit(86);
    }
    const long threads = static_cast<long>(block.x) * block.y * block.z;
    if (threads < 1 || threads > kMaxLanes || grid.x < 1 || grid.y < 1 || grid.z < 1) {
        std::fprintf(stderr, "emu: launch %s with an invalid shape\n", name);
        std::_Exit(86);
    }
    g.name = name;
    g.entry = entry;
    g.closure = closure;
    g.bdim = {block.x, block.y, block.z};
    g.gdim = {grid.x, grid.y, grid.z};
    g.n_lanes = static_cast<int>(threads);
    g.n_waves = (g.n_lanes + kWaveSize - 1) / kWaveSize;
    g.active = true;
    if (g_rng == 0) {
        g_rng = options().seed;
    }
    // The per-curve table kernels write their LDS before reading it and run
    // thousands of workgroups: only the sieve kernels' LDS is poisoned.
    const std::string kernel(name);
    g_poison_launch = options().poison && kernel.find("resid_kernel") == std::string::npos
        && kernel.find("chi_kernel") == std::string::npos;
    arm_watchdog(true);
    for (unsigned z = 0; z < grid.z; ++z) {
        for (unsigned y = 0; y < grid.y; ++y) {
            for (unsigned x = 0; x < grid.x; ++x) {
                g.block = {x, y, z};
                run_block();
            }
        }
    }
    arm_watchdog(false);
    g.active = false;
}

}  // namespace emu

// ---------------------------------------------------------------------------
// Host runtime API
// ---------------------------------------------------------------------------

struct EmuStream {
    unsigned flags = 0;
};

struct EmuEvent {
    std::chrono::steady_clock::time_point at{};
    bool recorded = false;
};

namespace {

int emu_cus() {
    const char *s = std::getenv("RPS_EMU_CUS");
    const int v = s ? std::atoi(s) : 4;
    return v > 0 ? v : 4;
}

}  // namespace

hipError_t hipGetDeviceCount(int *count) {
    *count = 1;
    return hipSuccess;
}

hipError_t hipSetDevice(int device) { return device == 0 ? hipSuccess : hipErrorInvalidDevice; }

hipError_t hipGetDevice(int *device) {
    *device = 0;
    return hipSuccess;
}

hipError_t hipGetDeviceProperties(hipDeviceProp_t *prop, int device) {
    if (device != 0) {
        return hipErrorInvalidDevice;
    }
    std::memset(prop, 0, sizeof *prop);
    std::snprintf(prop->name, sizeof prop->name, "CPU wave emulator");
    std::snprintf(prop->gcnArchName, sizeof prop->gcnArchName, "emu");
    prop->multiProcessorCount = emu_cus();
    prop->totalGlobalMem = size_t(1) << 34;
    prop->major = 9;
    prop->minor = 5;
    return hipSuccess;
}

hipError_t hipDeviceGetPCIBusId(char *bus, int len, int) {
    std::snprintf(bus, static_cast<size_t>(len), "emu");
    return hipSuccess;
}

const char *hipGetErrorString(hipError_t error) {
    switch (error) {
    case hipSuccess: return "no error";
    case hipErrorInvalidValue: return "invalid argument";
    case hipErrorOutOfMemory: return "out of memory";
    case hipErrorNotReady: return "device not ready";
    case hipErrorInvalidDevice: return "invalid device ordinal";
    default: return "unknown error";
    }
}

hipError_t hipGetLastError() { return hipSuccess; }

hipError_t hipMalloc(void **ptr, size_t bytes) {
    const size_t n = (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
    void *p = std::aligned_alloc(256, n);
    if (!p) {
        *ptr = nullptr;
        return hipErrorOutOfMemory;
    }
    std::memset(p, 0xCD, n);  // device memory starts undefined
    *ptr = p;
    return hipSuccess;
}

hipError_t hipFree(void *ptr) {
    std::free(ptr);
    return hipSuccess;
}

hipError_t hipHostMalloc(void **ptr, size_t bytes, unsigned) {
    return hipMalloc(ptr, bytes);
}

hipError_t hipHostFree(void *ptr) { return hipFree(ptr); }

hipError_t hipGetSymbolAddress(void **ptr, const void *symbol) {
    *ptr = const_cast<void *>(symbol);
    return hipSuccess;
}

hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind) {
    if (bytes) {
        std::memmove(dst, src, bytes);
    }
    return hipSuccess;
}

hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind,
                          hipStream_t) {
    return hipMemcpy(dst, src, bytes, kind);
}

hipError_t hipMemsetAsync(void *dst, int value, size_t bytes, hipStream_t) {
    if (bytes) {
        std::memset(dst, value, bytes);
    }
    return hipSuccess;
}

hipError_t hipMemset(void *dst, int value, size_t bytes) {
    return hipMemsetAsync(dst, value, bytes, nullptr);
}

hipError_t hipStreamCreate(hipStream_t *stream) {
    *stream = new EmuStream;
    return hipSuccess;
}

hipError_t hipStreamCreateWithFlags(hipStream_t *stream, unsigned flags) {
    *stream = new EmuStream;
    (*stream)->flags = flags;
    return hipSuccess;
}

hipError_t hipStreamDestroy(hipStream_t stream) {
    delete stream;
    return hipSuccess;
}

hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }

hipError_t hipEventCreate(hipEvent_t *event) {
    *event = new EmuEvent;
    return hipSuccess;
}

hipError_t hipEventCreateWithFlags(hipEvent_t *event, unsigned) { return hipEventCreate(event); }

hipError_t hipEventDestroy(hipEvent_t event) {
    delete event;
    return hipSuccess;
}

hipError_t hipEventRecord(hipEvent_t event, hipStream_t) {
    event->at = std::chrono::steady_clock::now();
    event->recorded = true;
    return hipSuccess;
}

hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }

hipError_t hipEventElapsedTime(float *ms, hipEvent_t start, hipEvent_t stop) {
    if (!start->recorded || !stop->recorded) {
        return hipErrorInvalidValue;
    }
    *ms = std::chrono::duration<float, std::milli>(stop->at - start->at).count();
    return hipSuccess;
}
