// ratpoints_gpu: bounded rational point search on y^2 = f(x) with the gfx950
// modular square sieve.
//
// The program's interface is the reference's (src/main.cpp): the invocations
// without a curve (no arguments, --help, --list-devices), exit status 1 with
// the message on stderr for any error, and the key=value lines of -v (or
// RATPOINTS_GPU_BENCHMARK) on stderr that tooling parses
// (tests/benchmark_multi_gpu.py in the reference).
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <string>
#include <string_view>

#include "command_line.hpp"
#include "gpu_devices.hpp"
#include "output_formatter.hpp"
#include "point_search.hpp"

namespace {

using namespace ratpoints_gpu;

// One line of space-separated key=value fields.
class FieldLine {
public:
    FieldLine &text(const char *key, const std::string &value) {
        if (!line_.empty()) {
            line_ += ' ';
        }
        line_ += key;
        line_ += '=';
        line_ += value;
        return *this;
    }
    FieldLine &count(const char *key, unsigned long long value) {
        return text(key, std::to_string(value));
    }
    FieldLine &ms(const char *key, double value) { return text(key, fixed3(value)); }

    void write(std::FILE *stream) const { std::fprintf(stream, "%s\n", line_.c_str()); }

private:
    static std::string fixed3(double v) {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%.3f", v);
        return buf;
    }

    std::string line_;
};

void write_device(std::FILE *stream, const GpuDevice &d) {
    FieldLine()
        .text("device", std::to_string(d.id))
        .text("name", "\"" + d.name + "\"")
        .text("compute", std::to_string(d.major) + "." + std::to_string(d.minor))
        .text("sm", std::to_string(d.multiprocessors))
        .count("memory_mib", d.memory_bytes / (1024ULL * 1024ULL))
        .text("pci", d.pci_bus_id)
        .write(stream);
}

void write_metrics(const SearchMetrics &m) {
    FieldLine()
        .ms("wall_ms", m.wall_ms)
        .count("devices", m.devices.size())
        .count("denominators", m.denominator_count)
        .ms("basis_ms", m.basis_ms)
        .ms("sieve_ms", m.sieve_ms)
        .count("words", m.word_count)
        .ms("initial_mask_gbs", m.initial_mask_gbs())
        .count("survivors", m.modular_survivors)
        .count("exact_survivors", m.exact_survivors)
        .ms("plan_ms", m.plan_ms)
        .ms("workspace_ms", m.workspace_ms)
        .ms("basis_wall_ms", m.basis_wall_ms)
        .ms("survivor_setup_ms", m.survivor_setup_ms)
        .ms("verification_ms", m.verification_ms)
        .ms("sieve_wall_ms", m.sieve_wall_ms)
        .write(stderr);
    for (const DeviceSearchMetrics &d : m.devices) {
        write_device(stderr, d.device);
        FieldLine()
            .text("gpu", std::to_string(d.device.id))
            .count("denominators", d.denominator_count)
            .count("batches", d.batches)
            .ms("basis_ms", d.basis_ms)
            .ms("sieve_ms", d.sieve_ms)
            .write(stderr);
    }
}

// A search: points to stdout as they are found, metrics afterwards.
int search(int argc, char **argv) {
    const ProgramOptions options = parse_command_line(argc, argv);
    PointSearch point_search(options.search);
    OutputFormatter output(options.output);
    const bool keep_going = !options.stop_after_first;
    const SearchMetrics metrics = point_search.run([&](const PointPair &point) {
        output.point_pair(point, keep_going);
        return keep_going;
    });
    output.finish();
    const bool wanted = std::getenv("RATPOINTS_GPU_BENCHMARK") != nullptr
        || (options.verbose && !options.quiet);
    if (wanted) {
        write_metrics(metrics);
    }
    return 0;
}

int list_devices() {
    for (const GpuDevice &d : available_gpu_devices()) {
        write_device(stdout, d);
    }
    return std::fflush(stdout) == EOF ? 1 : 0;
}

// Invocations without a curve: the whole command line is one word (or none).
struct Special {
    const char *word;  // nullptr: no arguments at all
    int (*run)();
};

const Special kSpecials[] = {
    {nullptr, [] { print_help(stderr); return 1; }},
    {"--help", [] { print_help(stdout); return 0; }},
    {"--list-devices", list_devices},
};

const Special *special_for(int argc, char **argv) {
    for (const Special &s : kSpecials) {
        const bool match = s.word ? argc == 2 && std::string_view(argv[1]) == s.word : argc == 1;
        if (match) {
            return &s;
        }
    }
    return nullptr;
}

}  // namespace

int main(int argc, char **argv) {
    try {
        const Special *special = special_for(argc, argv);
        return special ? special->run() : search(argc, argv);
    } catch (const std::exception &error) {
        std::fprintf(stderr, "%s\n", error.what());
        return 1;
    }
}
