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
// The Frobenius norm of a Krylov block (KY-E215 when it is not finite).
double block_scale_of(const Mat &block);
// g = r^T r (r upper triangular); false when a pivot is not positive or not finite.
bool cholesky_factor(const Mat &g, Mat &r);
// |r|_F |r^-1|_F, an upper bound of the 2-norm condition number of upper triangular r.
double condition_bound_of(const Mat &r);
Mat mul_nn(const Mat &a, const Mat &b);

// The basis B and its products C resident on the device (rows x cap each, row-major),
// for the dense steps' large FP64 products (dense_dev.hip). Holds the same columns as
// the host's basis and products.
struct DeviceBasis {
    Context *ctx;
    size_t rows, cap, cols = 0;
    DevBuf<double> basis, products, work_r, work_p, partial;
    DeviceBasis(Context &c, size_t rows, size_t cap);
    void append(const Mat &block, const Mat &product);
    // CGS2 against the held columns: rest -= B (B^T rest), twice.
    void project_out(Mat &rest);
    // The block on the device (work_r) with the held columns projected out, CGS2 as
    // project_out; it stays there for cholesky_qr2.
    void project_out_resident(const Mat &block);
    // CGS2 of the resident block (work_r, rows x width) against the held columns.
    void project_resident(size_t width);
    // Cholesky QR twice of the resident block (rows x w, w <= 64), in place: the
    // triangular factor into `r` and, with fetch_q, q into `q` (host); false when the
    // first factor's condition bound is above 1e6 or a pivot is not positive (then the
    // host's Householder QR takes the block, always downloaded into `q`).
    bool cholesky_qr2(size_t w, Mat &q, Mat &r, bool fetch_q = true);
    // The resident block (work_r, rows x width, contiguous): copied from device memory,
    // uploaded from the host, downloaded, or its Frobenius norm.
    void take_block(const double *src, size_t width);
    void put_block(const Mat &block);
    Mat get_block(size_t width);
    double block_norm(size_t width);
    // Columns [cols, cols + width) of B and C from device blocks (rows x width each,
    // contiguous), queued on the stream.
    void append_resident(const double *block, const double *product, size_t width);
    // B^T C (cols x cols).
    Mat rayleigh_matrix();
    // B S (rows x S.cols).
    Mat times(const Mat &small);
};

// pca_dense.rs:105-129. With `dev` (holding the same columns as `basis`), the
// projections run on the device.
Mat extend_basis(const Mat &basis, const Mat &block, DeviceBasis *dev = nullptr);
// The same over the columns `dev` holds (no host copy of the basis).
Mat extend_basis(DeviceBasis &dev, const Mat &block);
// The same for the block resident in dev.work_r (rows x width, Frobenius norm
// `scale`): the new directions are left there, contiguous; returns their number.
// They come back to the host only when the block needs the host's factorisations (more
// than 64 columns, an ill-conditioned block or a rank-deficient one).
size_t extend_basis_resident(DeviceBasis &dev, size_t width, double scale);

// pca_dense.rs:141-174. With `dev`, B^T C and B U run on the device.
void rayleigh_ritz(const Mat &basis, const Mat &products, size_t k, std::vector<double> &eigenvalues,
                   Mat &loadings, DeviceBasis *dev = nullptr);
// The same over the basis and products `dev` holds.
void rayleigh_ritz(DeviceBasis &dev, size_t k, std::vector<double> &eigenvalues, Mat &loadings);

}  // namespace ky
