// FP64 dense matrices and the PCA's small dense steps (host).
#pragma once

#include <cstddef>
#include <vector>

#include "common.hpp"

namespace ky {

// pca_dense.rs:30: a singular value at most this times the block's norm is dropped.
constexpr double DROP_TOLERANCE = 1e-10;

// A row-major rows x cols matrix of f64.
struct Mat {
    size_t rows = 0, cols = 0;
    std::vector<double> v;
    Mat() = default;
    Mat(size_t r, size_t c) : rows(r), cols(c), v(r * c, 0.0) {}
    double &at(size_t r, size_t c) { return v[r * cols + c]; }
    double at(size_t r, size_t c) const { return v[r * cols + c]; }
    double *row(size_t r) { return v.data() + r * cols; }
    const double *row(size_t r) const { return v.data() + r * cols; }
    Mat beside(const Mat &right) const;
};

Mat mul_tn(const Mat &a, const Mat &b);
Mat mul_nn(const Mat &a, const Mat &b);

// pca_dense.rs:105-129.
Mat extend_basis(const Mat &basis, const Mat &block);

// pca_dense.rs:141-174.
void rayleigh_ritz(const Mat &basis, const Mat &products, size_t k, std::vector<double> &eigenvalues,
                   Mat &loadings);

}  // namespace ky
