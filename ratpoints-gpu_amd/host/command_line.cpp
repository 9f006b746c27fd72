// Ratpoints-compatible command line.  Options, defaults, limits and error
// messages follow src/command_line.cpp of the reference (so scripts written
// for it keep working): ratpoints options -dl -du -l -u -1 -i -I -y -Y -z -Z
// -q -v -f -fs -fm -fe, the accepted no-ops -s -k -K -j -J, and the runtime
// options --devices --batch-size --square-denominators.
#include "command_line.hpp"

#include <cerrno>
#include <climits>
#include <cstdlib>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>

namespace ratpoints_gpu {
namespace {

constexpr size_t kMaxCoefficients = 101;  // degree 100
constexpr long long kMaxHeight = (LLONG_MAX - 32) / 2;
constexpr int kMaxSquareParameter = 46340;  // 46340^2 < 2^31
constexpr char kUsage[] = "Usage: ratpoints_gpu 'c0 c1 ... cn' HEIGHT [OPTIONS]";

long long positive(const char *text, const std::string &name) {
    errno = 0;
    char *end = nullptr;
    const long long value = std::strtoll(text, &end, 10);
    if (errno == ERANGE || end == text || *end != '\0' || value < 1) {
        throw std::invalid_argument(name + " must be positive");
    }
    return value;
}

int positive_int(const char *text, const std::string &name) {
    const long long value = positive(text, name);
    if (value > INT_MAX) {
        throw std::invalid_argument(name + " exceeds GPU limit");
    }
    return static_cast<int>(value);
}

// Decimal real with optional sign, fraction and exponent, as an exact
// rational: [+-]digits[.digits][(e|E)[+-]digits], |exponent| <= 10000.
mpq_class real(const char *text, const std::string &name) {
    const std::string s(text);
    size_t i = 0;
    bool negative = false;
    if (i < s.size() && (s[i] == '+' || s[i] == '-')) {
        negative = s[i++] == '-';
    }
    std::string digits;
    size_t fraction_digits = 0;
    bool any_digit = false;
    while (i < s.size() && s[i] >= '0' && s[i] <= '9') {
        digits += s[i++];
        any_digit = true;
    }
    if (i < s.size() && s[i] == '.') {
        ++i;
        const size_t mark = digits.size();
        while (i < s.size() && s[i] >= '0' && s[i] <= '9') {
            digits += s[i++];
            any_digit = true;
        }
        fraction_digits = digits.size() - mark;
    }
    long long exponent = 0;
    if (i < s.size() && (s[i] == 'e' || s[i] == 'E')) {
        ++i;
        bool exp_negative = false;
        if (i < s.size() && (s[i] == '+' || s[i] == '-')) {
            exp_negative = s[i++] == '-';
        }
        const size_t mark = i;
        while (i < s.size() && s[i] >= '0' && s[i] <= '9') {
            const int d = s[i++] - '0';
            if (exponent > (10000 - d) / 10) {
                throw std::invalid_argument(name + " exponent is too large");
            }
            exponent = 10 * exponent + d;
        }
        if (i == mark) {
            throw std::invalid_argument(name + " must be real");
        }
        if (exp_negative) {
            exponent = -exponent;
        }
    }
    if (!any_digit || i != s.size()) {
        throw std::invalid_argument(name + " must be real");
    }
    mpz_class mantissa(digits, 10);
    if (negative) {
        mantissa = -mantissa;
    }
    const long long scale = static_cast<long long>(fraction_digits) - exponent;
    mpz_class ten_power;
    mpz_ui_pow_ui(ten_power.get_mpz_t(), 10,
                  static_cast<unsigned long>(scale < 0 ? -scale : scale));
    mpq_class value = scale >= 0 ? mpq_class(mantissa, ten_power)
                                 : mpq_class(mantissa * ten_power);
    value.canonicalize();
    return value;
}

bool decimal_integer(const std::string &text) {
    size_t i = (!text.empty() && (text[0] == '+' || text[0] == '-')) ? 1 : 0;
    if (i == text.size()) {
        return false;
    }
    for (; i < text.size(); ++i) {
        if (text[i] < '0' || text[i] > '9') {
            return false;
        }
    }
    return true;
}

bool zero_text(const std::string &text) {
    size_t i = (!text.empty() && (text[0] == '+' || text[0] == '-')) ? 1 : 0;
    for (; i < text.size(); ++i) {
        if (text[i] != '0') {
            return false;
        }
    }
    return true;
}

Coefficients coefficient_list(const char *text) {
    std::istringstream in(text);
    Coefficients out;
    std::string token;
    while (in >> token) {
        if (token[0] == '+') {
            token.erase(0, 1);
        }
        out.push_back(std::move(token));
    }
    return out;
}

void check_coefficients(const Coefficients &coefficients) {
    if (coefficients.size() < 2 || coefficients.size() > kMaxCoefficients) {
        throw std::invalid_argument(
            "supported polynomial degrees are 1 through 100");
    }
    for (const std::string &c : coefficients) {
        if (!decimal_integer(c)) {
            throw std::invalid_argument("coefficients must be decimal integers");
        }
    }
    if (zero_text(coefficients.back())) {
        throw std::invalid_argument("leading coefficient must be nonzero");
    }
}

class Parser {
public:
    Parser(int argc, char **argv, ProgramOptions &out)
        : argc_(argc), argv_(argv), out_(out) {}

    void run() {
        while (next_ < argc_) {
            const std::string option = argv_[next_++];
            if (!search_option(option) && !output_option(option)
                && !runtime_option(option)) {
                throw std::invalid_argument("unsupported ratpoints option: "
                                            + option);
            }
        }
        if (open_lower_) {
            out_.search.intervals.push_back({lower_, mpq_class(), false, true});
        }
    }

private:
    const char *value_of(const std::string &option) {
        if (next_ >= argc_) {
            throw std::invalid_argument(option + " requires an argument");
        }
        return argv_[next_++];
    }

    bool search_option(const std::string &o) {
        if (o == "-dl") {
            out_.search.denominators.first =
                positive_int(value_of(o), "lower denominator");
        } else if (o == "-du") {
            out_.search.denominators.last =
                positive_int(value_of(o), "upper denominator");
        } else if (o == "-l") {
            if (open_lower_) {
                throw std::invalid_argument(
                    "interval lower bound without upper bound");
            }
            lower_ = real(value_of(o), "interval lower bound");
            open_lower_ = true;
        } else if (o == "-u") {
            close_interval(value_of(o));
        } else if (o == "-i") {
            out_.search.include_infinity = false;
        } else if (o == "-I") {
            out_.search.include_infinity = true;
        } else {
            return false;
        }
        return true;
    }

    bool output_option(const std::string &o) {
        OutputOptions &out = out_.output;
        if (o == "-y") {
            out.print_y = false;
        } else if (o == "-Y") {
            out.print_y = true;
        } else if (o == "-z") {
            out.print_points = false;
        } else if (o == "-Z") {
            out.print_points = true;
        } else if (o == "-f") {
            out.format = value_of(o);
        } else if (o == "-fs") {
            out.before = value_of(o);
        } else if (o == "-fm") {
            out.between = value_of(o);
        } else if (o == "-fe") {
            out.after = value_of(o);
        } else {
            return false;
        }
        return true;
    }

    bool runtime_option(const std::string &o) {
        if (o == "--devices") {
            out_.search.devices = parse_device_selection(value_of(o));
        } else if (o == "--batch-size") {
            out_.search.denominator_batch_size =
                positive_int(value_of(o), "denominator batch size");
            if (out_.search.denominator_batch_size > (1 << 16)) {
                throw std::invalid_argument("--batch-size must not exceed 65536");
            }
        } else if (o == "--square-denominators") {
            out_.search.square_denominators = true;
        } else if (o == "-q") {
            out_.quiet = true;
        } else if (o == "-v") {
            out_.verbose = true;
        } else if (o == "-1") {
            out_.stop_after_first = true;
        } else if (o == "-s" || o == "-k" || o == "-K" || o == "-j"
                   || o == "-J") {
            // Ratpoints CPU tuning switches; the searched set is unchanged.
        } else {
            return false;
        }
        return true;
    }

    void close_interval(const char *text) {
        const mpq_class upper = real(text, "interval upper bound");
        if (open_lower_ && lower_ > upper) {
            throw std::invalid_argument(
                "interval lower bound exceeds upper bound");
        }
        std::vector<SearchInterval> &list = out_.search.intervals;
        if (!list.empty()
            && (!open_lower_ || list.back().upper_infinite
                || lower_ < list.back().upper)) {
            throw std::invalid_argument("search intervals must be ordered");
        }
        list.push_back({lower_, upper, !open_lower_, false});
        open_lower_ = false;
    }

    int argc_;
    char **argv_;
    ProgramOptions &out_;
    int next_ = 3;
    mpq_class lower_;
    bool open_lower_ = false;
};

}  // namespace

bool help_requested(int argc, char **argv) {
    return argc == 2 && std::string(argv[1]) == "--help";
}

void print_help(std::FILE *stream) {
    static const char text[] =
        "Usage: ratpoints_gpu 'c0 c1 ... cn' HEIGHT [OPTIONS]\n"
        "\n"
        "GPU selection (IDs are relative to HIP_VISIBLE_DEVICES):\n"
        "  --devices all|N       select the GPU (default: the first visible)\n"
        "  --batch-size N        denominators per GPU batch, 1..65536\n"
        "  --square-denominators  interpret -dl/-du as k bounds; search d=k^2\n"
        "  --list-devices        list GPUs and exit; use without a curve\n"
        "\n"
        "Supported options:\n"
        "  -dl B       set the lower denominator bound\n"
        "  -du B       set the upper denominator bound\n"
        "  -l L        begin a closed x interval\n"
        "  -u U        end a closed x interval\n"
        "  -1          stop after the first point\n"
        "  -i, -I      suppress or restore points at infinity\n"
        "  -y, -Y      suppress or restore ordinates\n"
        "  -z, -Z      suppress or restore point output\n"
        "  -q          suppress timing output\n"
        "  -v          print timing and survivor metrics\n"
        "  -f FORMAT   set the point format\n"
        "  -fs TEXT    write TEXT before point output\n"
        "  -fm TEXT    write TEXT between points\n"
        "  -fe TEXT    write TEXT after point output\n"
        "  -s          accepted with no effect\n"
        "  -k, -K      accepted with no effect\n"
        "  -j, -J      accepted with no effect\n"
        "  --help      show this help\n";
    if (std::fputs(text, stream) == EOF) {
        throw std::runtime_error("failed to write help output");
    }
}

ProgramOptions parse_command_line(int argc, char **argv) {
    if (argc < 3) {
        throw std::invalid_argument(kUsage);
    }
    ProgramOptions out;
    out.search.coefficients = coefficient_list(argv[1]);
    out.search.numerator_bound = positive(argv[2], "maximum height");
    if (out.search.numerator_bound > kMaxHeight) {
        throw std::invalid_argument("maximum height exceeds safe GPU limit");
    }
    out.search.denominators.last = out.search.numerator_bound > INT_MAX
        ? INT_MAX : static_cast<int>(out.search.numerator_bound);
    Parser(argc, argv, out).run();
    if (out.search.denominators.last > out.search.numerator_bound) {
        out.search.denominators.last =
            static_cast<int>(out.search.numerator_bound);
    }
    if (out.search.square_denominators
        && out.search.denominators.last > kMaxSquareParameter) {
        out.search.denominators.last = kMaxSquareParameter;
    }
    if (out.search.denominators.first > out.search.denominators.last) {
        throw std::invalid_argument("empty denominator range");
    }
    check_coefficients(out.search.coefficients);
    return out;
}

}  // namespace ratpoints_gpu
