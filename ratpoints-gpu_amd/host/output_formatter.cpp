// Ratpoints-compatible point output.
//
// The bytes written are the reference's (src/output_formatter.cpp): the
// default formats "(%x : %y : %z)\n" and "(%x : %z)\n", the escapes \n, \t
// and \%, the -fs/-fm/-fe texts, the conjugate point (a : -y : b), and the
// -fe text written even under -z.  The design is this build's own: a format
// is compiled once into a list of literal pieces and fields, and each point
// is rendered from that list.
#include "output_formatter.hpp"

#include <cstdio>
#include <stdexcept>

namespace ratpoints_gpu {
namespace {

// Stands for an escaped '%' between decoding and field expansion, so that
// "\%x" stays the literal text "%x".
constexpr char kPercentMark = '\x01';

// Decodes \n, \t and \% (to kPercentMark); any other backslash, and a final
// one, is kept as written.
std::string decode_escapes(const std::string &text) {
    std::string out;
    out.reserve(text.size());
    std::size_t at = 0;
    while (at < text.size()) {
        const char c = text[at++];
        if (c != '\\' || at == text.size()) {
            out += c;
            continue;
        }
        const char e = text[at++];
        switch (e) {
        case 'n': out += '\n'; break;
        case 't': out += '\t'; break;
        case '%': out += kPercentMark; break;
        default:
            out += '\\';
            out += e;
        }
    }
    return out;
}

void put_text(const std::string &text) {
    if (std::fputs(text.c_str(), stdout) == EOF) {
        throw std::runtime_error("failed to write standard output");
    }
}

}  // namespace

OutputFormatter::OutputFormatter(const OutputOptions &options)
    : print_points_(options.print_points),
      print_y_(options.print_y),
      before_(decode_escapes(options.before)),
      between_(decode_escapes(options.between)),
      after_(decode_escapes(options.after)) {
    const std::string spec = decode_escapes(
        !options.format.empty() ? options.format
                                : print_y_ ? "(%x : %y : %z)\\n" : "(%x : %z)\\n");
    // %x, %z and (when ordinates are printed) %y become fields; any other
    // '%' and the character after it stay literal text, kPercentMark turns
    // back into '%'.
    std::string literal;
    auto flush_literal = [&]() {
        if (!literal.empty()) {
            for (char &c : literal) {
                if (c == kPercentMark) {
                    c = '%';
                }
            }
            format_.push_back({Piece::kText, literal});
            literal.clear();
        }
    };
    for (std::size_t at = 0; at < spec.size(); ++at) {
        if (spec[at] != '%' || at + 1 == spec.size()) {
            literal += spec[at];
            continue;
        }
        const char field = spec[++at];
        Piece::Kind kind = Piece::kText;
        if (field == 'x') {
            kind = Piece::kNumerator;
        } else if (field == 'z') {
            kind = Piece::kDenominator;
        } else if (field == 'y' && print_y_) {
            kind = Piece::kOrdinate;
        }
        if (kind == Piece::kText) {
            literal += '%';
            literal += field;
        } else {
            flush_literal();
            format_.push_back({kind, {}});
        }
    }
    flush_literal();
    if (print_points_) {
        put_text(before_);
    }
}

void OutputFormatter::point_pair(const PointPair &point, bool include_conjugate) {
    put_point(point.numerator, point.ordinate, point.denominator);
    const bool conjugate = include_conjugate && print_y_ && point.ordinate != "0";
    if (conjugate) {
        put_point(point.numerator, "-" + point.ordinate, point.denominator);
    }
}

void OutputFormatter::put_point(long long numerator, const std::string &ordinate,
                                int denominator) {
    if (!print_points_) {
        return;
    }
    if (any_point_) {
        put_text(between_);
    }
    std::string line;
    for (const Piece &piece : format_) {
        switch (piece.kind) {
        case Piece::kText: line += piece.text; break;
        case Piece::kNumerator: line += std::to_string(numerator); break;
        case Piece::kOrdinate: line += ordinate; break;
        case Piece::kDenominator: line += std::to_string(denominator); break;
        }
    }
    put_text(line);
    any_point_ = true;
}

void OutputFormatter::finish() {
    if (finished_) {
        return;
    }
    put_text(after_);  // also under -z
    if (std::fflush(stdout) == EOF) {
        throw std::runtime_error("failed to flush standard output");
    }
    finished_ = true;
}

}  // namespace ratpoints_gpu
