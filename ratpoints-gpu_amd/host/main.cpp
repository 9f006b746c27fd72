// ratpoints_gpu: bounded rational point search on y^2 = f(x) with the gfx950
// modular square sieve.  Command line, output and the -v metrics line follow
// src/main.cpp of the reference.
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <string>

#include "command_line.hpp"
#include "gpu_devices.hpp"
#include "output_formatter.hpp"
#include "point_search.hpp"

namespace {

using namespace ratpoints_gpu;

void describe(std::FILE *out, const GpuDevice &d) {
    std::fprintf(out,
                 "device=%d name=\"%s\" compute=%d.%d sm=%d memory_mib=%llu pci=%s\n",
                 d.id, d.name.c_str(), d.major, d.minor, d.multiprocessors,
                 d.memory_bytes / (1024ULL * 1024ULL), d.pci_bus_id.c_str());
}

// One line parsed by tooling as key=value pairs (src/main.cpp:22-43).
void report(const SearchMetrics &m) {
    std::fprintf(stderr,
                 "wall_ms=%.3f devices=%zu denominators=%llu "
                 "basis_ms=%.3f sieve_ms=%.3f words=%llu "
                 "initial_mask_gbs=%.3f survivors=%llu exact_survivors=%zu "
                 "plan_ms=%.3f workspace_ms=%.3f basis_wall_ms=%.3f "
                 "survivor_setup_ms=%.3f verification_ms=%.3f sieve_wall_ms=%.3f\n",
                 m.wall_ms, m.devices.size(), m.denominator_count, m.basis_ms,
                 m.sieve_ms, m.word_count, m.initial_mask_gbs(),
                 m.modular_survivors, m.exact_survivors, m.plan_ms,
                 m.workspace_ms, m.basis_wall_ms, m.survivor_setup_ms,
                 m.verification_ms, m.sieve_wall_ms);
    for (const DeviceSearchMetrics &d : m.devices) {
        describe(stderr, d.device);
        std::fprintf(stderr,
                     "gpu=%d denominators=%llu batches=%llu basis_ms=%.3f sieve_ms=%.3f\n",
                     d.device.id, d.denominator_count, d.batches, d.basis_ms,
                     d.sieve_ms);
    }
}

}  // namespace

int main(int argc, char **argv) {
    try {
        if (argc == 1) {
            print_help(stderr);
            return 1;
        }
        if (help_requested(argc, argv)) {
            print_help(stdout);
            return 0;
        }
        if (argc == 2 && std::string(argv[1]) == "--list-devices") {
            for (const GpuDevice &d : available_gpu_devices()) {
                describe(stdout, d);
            }
            return std::fflush(stdout) == EOF ? 1 : 0;
        }
        ProgramOptions options = parse_command_line(argc, argv);
        PointSearch search(options.search);
        OutputFormatter output(options.output);
        const bool first_only = options.stop_after_first;
        SearchMetrics metrics = search.run([&](const PointPair &point) {
            output.point_pair(point, !first_only);
            return !first_only;
        });
        output.finish();
        if (std::getenv("RATPOINTS_GPU_BENCHMARK") != nullptr
            || (options.verbose && !options.quiet)) {
            report(metrics);
        }
        return 0;
    } catch (const std::exception &error) {
        std::fprintf(stderr, "%s\n", error.what());
        return 1;
    }
}
