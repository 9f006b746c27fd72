// Exact stage: the homogenised value F(a, b) and its square root
// (src/exact_polynomial.cpp:25-42 of the reference), and the squarefree test
// gcd(f, f') = const (same file, 44-71).  The test first looks for a
// modular certificate -- a prime p not dividing the leading coefficient with
// gcd(f mod p, f' mod p) constant, which proves f squarefree over Q -- and
// runs Euclid's algorithm over Q only when no such prime turns up (in
// practice: when f has a repeated factor).  Over Q the remainder sequence of
// a degree-100 polynomial with 1,024-bit coefficients does not finish in
// minutes; the modular test takes microseconds.
#include "exact_polynomial.hpp"

#include <cstdint>
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

uint64_t pow_mod(uint64_t base, uint64_t e, uint64_t p) {
    uint64_t r = 1;
    base %= p;
    while (e) {
        if (e & 1) {
            r = r * base % p;
        }
        base = base * base % p;
        e >>= 1;
    }
    return r;
}

void trim_mod(std::vector<uint64_t> &a) {
    while (!a.empty() && a.back() == 0) {
        a.pop_back();
    }
}

// Degree of gcd(a, b) over F_p (p < 2^32 prime, so products fit in 64
// bits); -1 when both are zero.
int gcd_degree_mod(std::vector<uint64_t> a, std::vector<uint64_t> b, uint64_t p) {
    trim_mod(a);
    trim_mod(b);
    while (!b.empty()) {
        const uint64_t inv = pow_mod(b.back(), p - 2, p);
        while (a.size() >= b.size()) {
            const size_t shift = a.size() - b.size();
            const uint64_t q = a.back() * inv % p;
            for (size_t i = 0; i < b.size(); ++i) {
                a[shift + i] = (a[shift + i] + (p - q) * b[i]) % p;
            }
            trim_mod(a);
        }
        std::swap(a, b);
    }
    return static_cast<int>(a.size()) - 1;
}

// A prime p with p not dividing lc(f) and gcd(f mod p, f' mod p) constant
// proves f squarefree: a repeated factor g of f over Q can be taken
// primitive in Z[x] (Gauss), then g^2 divides f in Z[x], deg(g mod p) =
// deg g because p does not divide lc(f), and g mod p divides f mod p and
// f' mod p.
bool modular_squarefree_certificate(const std::vector<mpz_class> &f) {
    static const uint32_t kPrimes[] = {2147483647u, 2147483629u, 2147483587u,
                                       2147483579u, 2147483563u, 2147483549u};
    for (uint32_t prime : kPrimes) {
        const uint64_t p = prime;
        std::vector<uint64_t> a(f.size()), da(f.size() - 1);
        for (size_t i = 0; i < f.size(); ++i) {
            a[i] = mpz_fdiv_ui(f[i].get_mpz_t(), prime);
        }
        if (a.back() == 0) {
            continue;
        }
        for (size_t i = 1; i < f.size(); ++i) {
            da[i - 1] = a[i] * (i % p) % p;
        }
        if (gcd_degree_mod(a, da, p) == 0) {
            return true;
        }
    }
    return false;
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
    std::vector<mpz_class> z;
    z.reserve(coefficients.size());
    for (const auto &text : coefficients) {
        z.emplace_back(text, 10);
    }
    while (!z.empty() && sgn(z.back()) == 0) {
        z.pop_back();
    }
    if (z.size() >= 2 && modular_squarefree_certificate(z)) {
        return true;
    }
    QPoly f(z.begin(), z.end());
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
