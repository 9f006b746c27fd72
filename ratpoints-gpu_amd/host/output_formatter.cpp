// Point output with ratpoints' format markers %x %y %z and the escapes \n \t
// \% (behaviour of src/output_formatter.cpp in the reference).
#include "output_formatter.hpp"

#include <cstdio>
#include <stdexcept>

namespace ratpoints_gpu {
namespace {

// Placeholder for an escaped percent sign until markers are expanded.
constexpr char kLiteralPercent = '\x01';

}  // namespace

OutputFormatter::OutputFormatter(const OutputOptions &options)
    : print_points_(options.print_points),
      print_y_(options.print_y),
      format_(unescape(!options.format.empty() ? options.format
                       : options.print_y ? std::string("(%x : %y : %z)\\n")
                                         : std::string("(%x : %z)\\n"))),
      before_(unescape(options.before)),
      between_(unescape(options.between)),
      after_(unescape(options.after)) {
    if (print_points_) {
        emit(before_);
    }
}

void OutputFormatter::point_pair(const PointPair &point,
                                 bool include_conjugate) {
    write_point(point.numerator, point.ordinate, point.denominator);
    if (include_conjugate && print_y_ && point.ordinate != "0") {
        write_point(point.numerator, "-" + point.ordinate, point.denominator);
    }
}

void OutputFormatter::write_point(long long numerator,
                                  const std::string &ordinate,
                                  int denominator) {
    if (!print_points_) {
        return;
    }
    if (wrote_point_) {
        emit(between_);
    }
    std::string line;
    line.reserve(format_.size() + ordinate.size() + 40);
    const size_t n = format_.size();
    for (size_t i = 0; i < n; ++i) {
        const char ch = format_[i];
        if (ch != '%' || i + 1 == n) {
            line += ch == kLiteralPercent ? '%' : ch;
            continue;
        }
        const char marker = format_[++i];
        switch (marker) {
        case 'x':
            line += std::to_string(numerator);
            break;
        case 'y':
            // With ordinates suppressed, %y is kept as literal text.
            if (print_y_) {
                line += ordinate;
            } else {
                line += "%y";
            }
            break;
        case 'z':
            line += std::to_string(denominator);
            break;
        default:
            line += '%';
            line += marker == kLiteralPercent ? '%' : marker;
        }
    }
    emit(line);
    wrote_point_ = true;
}

void OutputFormatter::finish() {
    if (finished_) {
        return;
    }
    emit(after_);
    if (std::fflush(stdout) == EOF) {
        throw std::runtime_error("failed to flush standard output");
    }
    finished_ = true;
}

std::string OutputFormatter::unescape(const std::string &text) {
    std::string out;
    out.reserve(text.size());
    for (size_t i = 0; i < text.size(); ++i) {
        if (text[i] != '\\' || i + 1 == text.size()) {
            out += text[i];
            continue;
        }
        const char next = text[++i];
        if (next == 'n') {
            out += '\n';
        } else if (next == 't') {
            out += '\t';
        } else if (next == '%') {
            out += kLiteralPercent;
        } else {
            out += '\\';
            out += next;
        }
    }
    return out;
}

void OutputFormatter::emit(const std::string &text) {
    if (std::fputs(text.c_str(), stdout) == EOF) {
        throw std::runtime_error("failed to write standard output");
    }
}

}  // namespace ratpoints_gpu
