// Ratpoints-compatible command line.
//
// The interface is the reference's (src/command_line.cpp): the options, their
// defaults and limits, the error messages and the order in which they are
// checked, so scripts written for ratpoints_gpu keep working.  The parser is
// this build's own: options are a table of rules, each rule pulls the
// arguments it needs from a cursor over argv, decimal reals are split into
// their parts by a tokenizer before any value is formed, and the search
// options are checked as a whole once every option has been read.
#include "command_line.hpp"

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <climits>
#include <cstdlib>
#include <optional>
#include <stdexcept>
#include <string>
#include <string_view>
#include <utility>

namespace ratpoints_gpu {
namespace {

constexpr std::size_t kMostTerms = 101;              // degree 1 .. 100
constexpr long long kHeightLimit = (LLONG_MAX - 32) / 2;
constexpr int kLargestSquareParameter = 46340;       // 46340^2 < 2^31
constexpr long long kLargestBatch = 1 << 16;
constexpr long kExponentLimit = 10000;

[[noreturn]] void reject(const std::string &message) {
    throw std::invalid_argument(message);
}

// ---------------------------------------------------------------------------
// Numbers
// ---------------------------------------------------------------------------

// A whole argument as a positive long long (strtoll rules: leading blanks and
// a sign are accepted, nothing may follow the digits).
long long positive_value(const char *text, const std::string &what) {
    char *rest = nullptr;
    errno = 0;
    const long long v = std::strtoll(text, &rest, 10);
    const bool whole = rest != text && *rest == '\0';
    if (!whole || errno == ERANGE || v <= 0) {
        reject(what + " must be positive");
    }
    return v;
}

int positive_int_value(const char *text, const std::string &what) {
    const long long v = positive_value(text, what);
    if (v > INT_MAX) {
        reject(what + " exceeds GPU limit");
    }
    return static_cast<int>(v);
}

bool is_digit(char c) { return c >= '0' && c <= '9'; }

// The parts of [+-]digits[.digits][(e|E)[+-]digits] found at the start of a
// text; `matched` is how many characters they cover.
struct DecimalParts {
    bool negative = false;
    std::string digits;             // integer digits then fraction digits
    std::size_t fraction_digits = 0;
    bool has_exponent = false;
    bool exponent_negative = false;
    std::string_view exponent_digits;
    std::size_t matched = 0;
};

std::size_t end_of_digits(std::string_view s, std::size_t from) {
    while (from < s.size() && is_digit(s[from])) {
        ++from;
    }
    return from;
}

DecimalParts split_decimal(std::string_view s) {
    DecimalParts d;
    std::size_t at = 0;
    if (at < s.size() && (s[at] == '+' || s[at] == '-')) {
        d.negative = s[at] == '-';
        ++at;
    }
    std::size_t stop = end_of_digits(s, at);
    d.digits.assign(s.substr(at, stop - at));
    at = stop;
    if (at < s.size() && s[at] == '.') {
        stop = end_of_digits(s, at + 1);
        d.fraction_digits = stop - at - 1;
        d.digits.append(s.substr(at + 1, d.fraction_digits));
        at = stop;
    }
    if (at < s.size() && (s[at] == 'e' || s[at] == 'E')) {
        d.has_exponent = true;
        ++at;
        if (at < s.size() && (s[at] == '+' || s[at] == '-')) {
            d.exponent_negative = s[at] == '-';
            ++at;
        }
        stop = end_of_digits(s, at);
        d.exponent_digits = s.substr(at, stop - at);
        at = stop;
    }
    d.matched = at;
    return d;
}

// An interval end given as a decimal real, as an exact rational.  Exponents
// are limited to 10,000 in magnitude.
mpq_class exact_decimal(const char *text, const std::string &what) {
    const std::string_view s(text);
    const DecimalParts d = split_decimal(s);
    long exponent = 0;
    for (char c : d.exponent_digits) {
        exponent = 10 * exponent + (c - '0');
        if (exponent > kExponentLimit) {
            reject(what + " exponent is too large");
        }
    }
    const bool exponent_ok = !d.has_exponent || !d.exponent_digits.empty();
    if (!exponent_ok || d.digits.empty() || d.matched != s.size()) {
        reject(what + " must be real");
    }
    mpz_class mantissa(d.digits, 10);
    if (d.negative) {
        mantissa = -mantissa;
    }
    // value = mantissa * 10^(exponent - fraction digits)
    const long shift = (d.exponent_negative ? -exponent : exponent)
        - static_cast<long>(d.fraction_digits);
    mpz_class ten_power;
    mpz_ui_pow_ui(ten_power.get_mpz_t(), 10, static_cast<unsigned long>(shift < 0 ? -shift : shift));
    mpq_class value = shift < 0 ? mpq_class(mantissa, ten_power) : mpq_class(mantissa * ten_power);
    value.canonicalize();
    return value;
}

// ---------------------------------------------------------------------------
// Coefficients
// ---------------------------------------------------------------------------

// Blank-separated terms c0 .. cn; one leading '+' of a term is dropped.
Coefficients split_terms(const char *text) {
    Coefficients terms;
    std::string current;
    auto close = [&]() {
        if (!current.empty()) {
            terms.push_back(current.front() == '+' ? current.substr(1) : current);
            current.clear();
        }
    };
    for (const char *p = text; *p; ++p) {
        if (std::isspace(static_cast<unsigned char>(*p))) {
            close();
        } else {
            current.push_back(*p);
        }
    }
    close();
    return terms;
}

// [+-]digits, at least one digit.
bool integer_text(const std::string &t) {
    const std::size_t from = !t.empty() && (t[0] == '+' || t[0] == '-') ? 1 : 0;
    return from < t.size() && end_of_digits(t, from) == t.size();
}

bool zero_text(const std::string &t) {
    const std::size_t from = !t.empty() && (t[0] == '+' || t[0] == '-') ? 1 : 0;
    return t.find_first_not_of('0', from) == std::string::npos;
}

void check_terms(const Coefficients &terms) {
    if (terms.size() < 2 || terms.size() > kMostTerms) {
        reject("supported polynomial degrees are 1 through 100");
    }
    for (const std::string &t : terms) {
        if (!integer_text(t)) {
            reject("coefficients must be decimal integers");
        }
    }
    if (zero_text(terms.back())) {
        reject("leading coefficient must be nonzero");
    }
}

// ---------------------------------------------------------------------------
// Options
// ---------------------------------------------------------------------------

// argv after the curve and the height.
class ArgCursor {
public:
    ArgCursor(int argc, char **argv) : end_(argv + argc), next_(argv + 3) {}

    bool more() const { return next_ < end_; }
    std::string option() { return std::string(*next_++); }

    // The argument of `option`; a missing one is an error.
    const char *value_of(const std::string &option) {
        if (next_ >= end_) {
            reject(option + " requires an argument");
        }
        return *next_++;
    }

private:
    char **end_;
    char **next_;
};

// Everything the options write: the program options and the lower end of
// an interval that -l opened and no -u has closed yet.
struct ParseState {
    ProgramOptions &out;
    std::optional<mpq_class> open_lower;
};

using Rule = void (*)(ParseState &, ArgCursor &, const std::string &);

struct OptionRule {
    const char *name;
    Rule apply;
};

void open_interval(ParseState &st, ArgCursor &args, const std::string &name) {
    if (st.open_lower) {
        reject("interval lower bound without upper bound");
    }
    st.open_lower = exact_decimal(args.value_of(name), "interval lower bound");
}

void close_interval(ParseState &st, ArgCursor &args, const std::string &name) {
    const mpq_class upper = exact_decimal(args.value_of(name), "interval upper bound");
    if (st.open_lower && *st.open_lower > upper) {
        reject("interval lower bound exceeds upper bound");
    }
    std::vector<SearchInterval> &list = st.out.search.intervals;
    if (!list.empty()) {
        const SearchInterval &last = list.back();
        const bool follows = st.open_lower && !last.upper_infinite
            && !(*st.open_lower < last.upper);
        if (!follows) {
            reject("search intervals must be ordered");
        }
    }
    SearchInterval added;
    added.lower_infinite = !st.open_lower;
    if (st.open_lower) {
        added.lower = *st.open_lower;
    }
    added.upper = upper;
    list.push_back(added);
    st.open_lower.reset();
}

const OptionRule kRules[] = {
    // search box
    {"-dl", [](ParseState &st, ArgCursor &a, const std::string &o) {
         st.out.search.denominators.first = positive_int_value(a.value_of(o), "lower denominator");
     }},
    {"-du", [](ParseState &st, ArgCursor &a, const std::string &o) {
         st.out.search.denominators.last = positive_int_value(a.value_of(o), "upper denominator");
     }},
    {"-l", open_interval},
    {"-u", close_interval},
    {"-i", [](ParseState &st, ArgCursor &, const std::string &) {
         st.out.search.include_infinity = false;
     }},
    {"-I", [](ParseState &st, ArgCursor &, const std::string &) {
         st.out.search.include_infinity = true;
     }},
    // output
    {"-y", [](ParseState &st, ArgCursor &, const std::string &) { st.out.output.print_y = false; }},
    {"-Y", [](ParseState &st, ArgCursor &, const std::string &) { st.out.output.print_y = true; }},
    {"-z", [](ParseState &st, ArgCursor &, const std::string &) {
         st.out.output.print_points = false;
     }},
    {"-Z", [](ParseState &st, ArgCursor &, const std::string &) {
         st.out.output.print_points = true;
     }},
    {"-f", [](ParseState &st, ArgCursor &a, const std::string &o) {
         st.out.output.format = a.value_of(o);
     }},
    {"-fs", [](ParseState &st, ArgCursor &a, const std::string &o) {
         st.out.output.before = a.value_of(o);
     }},
    {"-fm", [](ParseState &st, ArgCursor &a, const std::string &o) {
         st.out.output.between = a.value_of(o);
     }},
    {"-fe", [](ParseState &st, ArgCursor &a, const std::string &o) {
         st.out.output.after = a.value_of(o);
     }},
    // run control
    {"--devices", [](ParseState &st, ArgCursor &a, const std::string &o) {
         st.out.search.devices = parse_device_selection(a.value_of(o));
     }},
    {"--batch-size", [](ParseState &st, ArgCursor &a, const std::string &o) {
         const int n = positive_int_value(a.value_of(o), "denominator batch size");
         if (n > kLargestBatch) {
             reject("--batch-size must not exceed 65536");
         }
         st.out.search.denominator_batch_size = n;
     }},
    {"--square-denominators", [](ParseState &st, ArgCursor &, const std::string &) {
         st.out.search.square_denominators = true;
     }},
    {"-q", [](ParseState &st, ArgCursor &, const std::string &) { st.out.quiet = true; }},
    {"-v", [](ParseState &st, ArgCursor &, const std::string &) { st.out.verbose = true; }},
    {"-1", [](ParseState &st, ArgCursor &, const std::string &) { st.out.stop_after_first = true; }},
    // ratpoints' CPU tuning switches: accepted, the searched set is the same
    {"-s", [](ParseState &, ArgCursor &, const std::string &) {}},
    {"-k", [](ParseState &, ArgCursor &, const std::string &) {}},
    {"-K", [](ParseState &, ArgCursor &, const std::string &) {}},
    {"-j", [](ParseState &, ArgCursor &, const std::string &) {}},
    {"-J", [](ParseState &, ArgCursor &, const std::string &) {}},
};

const OptionRule *find_rule(const std::string &name) {
    for (const OptionRule &r : kRules) {
        if (name == r.name) {
            return &r;
        }
    }
    return nullptr;
}

void read_options(int argc, char **argv, ProgramOptions &out) {
    ParseState st{out, std::nullopt};
    ArgCursor args(argc, argv);
    while (args.more()) {
        const std::string name = args.option();
        const OptionRule *rule = find_rule(name);
        if (!rule) {
            reject("unsupported ratpoints option: " + name);
        }
        rule->apply(st, args, name);
    }
    if (st.open_lower) {
        // -l without -u: the interval is [lower, +inf).
        SearchInterval tail;
        tail.lower = *st.open_lower;
        tail.upper_infinite = true;
        out.search.intervals.push_back(tail);
    }
}

// The denominator range after every option: -du is capped at the height,
// and at 46,340 in square mode (k^2 must stay below 2^31).
void settle_denominators(SearchOptions &search) {
    DenominatorRange &range = search.denominators;
    if (range.last > search.numerator_bound) {
        range.last = static_cast<int>(search.numerator_bound);
    }
    if (search.square_denominators && range.last > kLargestSquareParameter) {
        range.last = kLargestSquareParameter;
    }
    if (range.first > range.last) {
        reject("empty denominator range");
    }
}

const char kHelpText[] =
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

}  // namespace

bool help_requested(int argc, char **argv) {
    return argc == 2 && std::string_view(argv[1]) == "--help";
}

void print_help(std::FILE *stream) {
    if (std::fputs(kHelpText, stream) == EOF) {
        throw std::runtime_error("failed to write help output");
    }
}

ProgramOptions parse_command_line(int argc, char **argv) {
    if (argc < 3) {
        reject("Usage: ratpoints_gpu 'c0 c1 ... cn' HEIGHT [OPTIONS]");
    }
    ProgramOptions out;
    SearchOptions &search = out.search;
    search.coefficients = split_terms(argv[1]);
    search.numerator_bound = positive_value(argv[2], "maximum height");
    if (search.numerator_bound > kHeightLimit) {
        reject("maximum height exceeds safe GPU limit");
    }
    // -du defaults to the height.
    search.denominators.last = static_cast<int>(std::min<long long>(search.numerator_bound, INT_MAX));
    read_options(argc, argv, out);
    settle_denominators(search);
    check_terms(search.coefficients);
    return out;
}

}  // namespace ratpoints_gpu
