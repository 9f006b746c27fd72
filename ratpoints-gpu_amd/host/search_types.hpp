// Value types shared by the search layers.  Same names and meaning as the
// reference's include/search_types.hpp:9-22, so callers written against the
// reference compile unchanged.
#ifndef RATPOINTS_GPU_SEARCH_TYPES_HPP
#define RATPOINTS_GPU_SEARCH_TYPES_HPP

#include <string>
#include <vector>

namespace ratpoints_gpu {

// Decimal coefficients c0 .. cn of f(x) = c0 + c1 x + ... + cn x^n.
using Coefficients = std::vector<std::string>;

// Inclusive range of denominator parameters (b, or k with b = k^2).
struct DenominatorRange {
    int first;
    int last;

    int count() const { return last - first + 1; }
};

// A verified point (numerator : ordinate : denominator); ordinate >= 0.
struct PointPair {
    long long numerator;
    std::string ordinate;
    int denominator;
};

}  // namespace ratpoints_gpu

#endif
