// Exact stage: the homogenised value F(a, b) and its square root
// (src/exact_polynomial.cpp:25-42 of the reference), and the squarefree test
// gcd(f, f') = const by Euclid's algorithm over Q (same file, 44-71).
#include "exact_polynomial.hpp"

#include <utility>

namespace ratpoints_gpu {
namespace {

using QPoly = std::vector<mpq_class>;

void drop_leading_zeros(QPoly &p) {
    while (!p.empty() && sgn(p.back()) == 0) {
        p.pop_back();
    }
}

// Remainder of a by b (b nonzero with a nonzero leading coefficient).
QPoly remainder(QPoly a, const QPoly &b) {
    drop_leading_zeros(a);
    while (a.size() >= b.size()) {
        const size_t shift = a.size() - b.size();
        const mpq_class q = a.back() / b.back();
        for (size_t i = 0; i < b.size(); ++i) {
            a[shift + i] -= q * b[i];
        }
        a.pop_back();  // the leading term cancels exactly
        drop_leading_zeros(a);
    }
    return a;
}

struct Scratch {
    mpz_class value;
    mpz_class power;
    mpz_class term;
};

thread_local Scratch scratch;

}  // namespace

ExactPolynomial::ExactPolynomial(const Coefficients &coefficients) {
    coefficients_.reserve(coefficients.size());
    for (const auto &text : coefficients) {
        coefficients_.emplace_back(text, 10);
    }
}

bool ExactPolynomial::square_root(long long numerator, int denominator,
                                  mpz_class &root) const {
    // Horner on the binary form: v <- v a + c_i b^(n-i), from the top.
    Scratch &s = scratch;
    const size_t n = coefficients_.size() - 1;
    s.value = coefficients_[n];
    s.power = 1;
    for (size_t k = n; k-- > 0;) {
        mpz_mul_si(s.value.get_mpz_t(), s.value.get_mpz_t(), numerator);
        mpz_mul_si(s.power.get_mpz_t(), s.power.get_mpz_t(), denominator);
        mpz_mul(s.term.get_mpz_t(), coefficients_[k].get_mpz_t(),
                s.power.get_mpz_t());
        s.value += s.term;
    }
    if (n % 2 == 1) {
        // Odd degree: F(a, b) = b * b^n f(a / b) keeps y^2 = F integral.
        mpz_mul_si(s.value.get_mpz_t(), s.value.get_mpz_t(), denominator);
    }
    if (sgn(s.value) < 0 || !mpz_perfect_square_p(s.value.get_mpz_t())) {
        return false;
    }
    mpz_sqrt(root.get_mpz_t(), s.value.get_mpz_t());
    return true;
}

bool ExactPolynomial::is_squarefree(const Coefficients &coefficients) {
    QPoly f;
    f.reserve(coefficients.size());
    for (const auto &text : coefficients) {
        f.emplace_back(mpz_class(text, 10));
    }
    drop_leading_zeros(f);
    QPoly df;
    for (size_t i = 1; i < f.size(); ++i) {
        df.push_back(f[i] * static_cast<unsigned long>(i));
    }
    drop_leading_zeros(df);
    // gcd(f, f') by Euclid; f is squarefree iff the gcd is a constant.
    while (!df.empty()) {
        QPoly r = remainder(f, df);
        f = std::move(df);
        df = std::move(r);
    }
    return f.size() == 1;
}

}  // namespace ratpoints_gpu
