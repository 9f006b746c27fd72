// Exact arithmetic on f with GMP (the reference's
// include/exact_polynomial.hpp:11-21, same interface).
#ifndef RATPOINTS_GPU_EXACT_POLYNOMIAL_HPP
#define RATPOINTS_GPU_EXACT_POLYNOMIAL_HPP

#include <gmpxx.h>

#include <vector>

#include "search_types.hpp"

namespace ratpoints_gpu {

class ExactPolynomial {
public:
    explicit ExactPolynomial(const Coefficients &coefficients);

    // root = sqrt(F(a, b)) when F(a, b) = sum c_i a^i b^(n-i) (times b for
    // odd n) is a nonnegative perfect square; false otherwise.
    bool square_root(long long numerator, int denominator,
                     mpz_class &root) const;
    // gcd(f, f') is a nonzero constant over Q.
    static bool is_squarefree(const Coefficients &coefficients);

private:
    std::vector<mpz_class> coefficients_;
};

}  // namespace ratpoints_gpu

#endif
