// Ratpoints-compatible point output (the reference's
// include/output_formatter.hpp:9-45, same interface and behaviour).
#ifndef RATPOINTS_GPU_OUTPUT_FORMATTER_HPP
#define RATPOINTS_GPU_OUTPUT_FORMATTER_HPP

#include <string>
#include <vector>

#include "search_types.hpp"

namespace ratpoints_gpu {

struct OutputOptions {
    bool print_points = true;  // -z / -Z
    bool print_y = true;       // -y / -Y
    std::string format;        // -f; empty selects the default format
    std::string before;        // -fs
    std::string between;       // -fm
    std::string after;         // -fe
};

class OutputFormatter {
public:
    explicit OutputFormatter(const OutputOptions &options);

    OutputFormatter(const OutputFormatter &) = delete;
    OutputFormatter &operator=(const OutputFormatter &) = delete;

    // Writes the point, then (a : -y : b) when y != 0, ordinates are printed
    // and include_conjugate is set.
    void point_pair(const PointPair &point, bool include_conjugate = true);
    // Writes the -fe text (also under -z) and flushes stdout, once.
    void finish();

private:
    // One piece of a compiled point format: literal text or a field.
    struct Piece {
        enum Kind { kText, kNumerator, kOrdinate, kDenominator } kind;
        std::string text;
    };

    void put_point(long long numerator, const std::string &ordinate, int denominator);

    bool print_points_;
    bool print_y_;
    std::vector<Piece> format_;
    std::string before_;
    std::string between_;
    std::string after_;
    bool any_point_ = false;
    bool finished_ = false;
};

}  // namespace ratpoints_gpu

#endif
