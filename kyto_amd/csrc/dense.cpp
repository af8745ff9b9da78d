// The FP64 dense steps of the streamed PCA on the host: kyto's extend_basis and
// rayleigh_ritz (crates/kyto-exec/src/pca_dense.rs:105-273), restated over this
// library's own FP64 routines (kyto uses faer 0.24.4 for these; its results are
// pinned by tolerances, pca_dense.rs:285-355):
// - thin SVD: Householder QR, then one-sided Jacobi (Hestenes) on the small R;
// - symmetric eigenpairs: Householder tridiagonalisation and the implicit QL method
//   with Wilkinson-style shifts (the EISPACK tred2/tql2 pair).
// Products, reflections and rotations are cut over a persistent pool of worker
// threads; each output element is computed in one fixed order, so the bits do not
// depend on the thread count.
#include "dense.hpp"

#include <algorithm>
#include <cstdint>
#include <cmath>
#include <numeric>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>

namespace ky {

// A persistent pool of worker threads for the dense steps. run(tasks, f) calls f(t)
// for every t in [0, tasks), task t on worker t % size() (the caller is worker 0), and
// returns when every task is done; each output element is written by one task in a
// fixed order, so the bits do not depend on the number of workers. Workers spin for a
// short while between runs (the steps of a factorisation come microseconds apart),
// then sleep. One run at a time (a mutex serialises callers).
class Pool {
  public:
    static Pool &get() {
        static Pool pool;
        return pool;
    }
    unsigned size() const { return workers_ + 1; }
    // Threads a run uses (the caller and active - 1 workers; at most size()).
    void set_active(unsigned threads) { active_.store(threads == 0 ? size() : std::min(threads, size())); }
    unsigned active() const { return active_.load(); }
    template <typename F>
    void run(size_t tasks, F &&f) {
        if (tasks == 0) return;
        const unsigned active = active_.load();
        if (active <= 1 || tasks == 1) {
            for (size_t t = 0; t < tasks; t++) f(t);
            return;
        }
        std::lock_guard<std::mutex> serial(run_mutex_);
        job_ = [&f](size_t t) { f(t); };
        tasks_ = tasks;
        stride_ = active;
        done_.store(0, std::memory_order_relaxed);
        {
            std::lock_guard<std::mutex> lock(sleep_mutex_);
            epoch_.fetch_add(1, std::memory_order_release);
        }
        sleep_cv_.notify_all();
        for (size_t t = 0; t < tasks; t += stride_) job_(t);
        while (done_.load(std::memory_order_acquire) != workers_) spin_pause();
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> lock(sleep_mutex_);
            stop_ = true;
            epoch_.fetch_add(1, std::memory_order_release);
        }
        sleep_cv_.notify_all();
        for (auto &t : threads_) t.join();
    }

  private:
    Pool() {
        unsigned n = std::thread::hardware_concurrency();
        workers_ = std::max(1u, std::min(n, 32u)) - 1;
        active_ = workers_ + 1;
        for (unsigned w = 1; w <= workers_; w++) threads_.emplace_back([this, w] { loop(w); });
    }
    static void spin_pause() { __builtin_ia32_pause(); }
    void loop(unsigned w) {
        uint64_t seen = 0;
        for (;;) {
            uint64_t e = epoch_.load(std::memory_order_acquire);
            for (int spin = 0; e == seen && spin < 200000; spin++) {
                spin_pause();
                e = epoch_.load(std::memory_order_acquire);
            }
            if (e == seen) {
                std::unique_lock<std::mutex> lock(sleep_mutex_);
                sleep_cv_.wait(lock, [&] { return epoch_.load(std::memory_order_acquire) != seen; });
                e = epoch_.load(std::memory_order_acquire);
            }
            seen = e;
            if (stop_) return;
            if (w < stride_)
                for (size_t t = w; t < tasks_; t += stride_) job_(t);
            done_.fetch_add(1, std::memory_order_acq_rel);
        }
    }
    unsigned workers_ = 0, stride_ = 1;
    std::atomic<unsigned> active_{1};
    std::vector<std::thread> threads_;
    std::function<void(size_t)> job_;
    size_t tasks_ = 0;
    std::atomic<uint64_t> epoch_{0};
    std::atomic<unsigned> done_{0};
    bool stop_ = false;
    std::mutex run_mutex_, sleep_mutex_;
    std::condition_variable sleep_cv_;
};

void host_parallel(size_t tasks, const std::function<void(size_t)> &f) { Pool::get().run(tasks, f); }

void host_threads(unsigned threads) { Pool::get().set_active(threads); }
unsigned host_threads_active() { return Pool::get().active(); }

// f(begin, end) over [0, n) in contiguous ranges, one a worker (serially below 64).
template <typename F>
static void parallel_for(size_t n, F &&f) {
    Pool &pool = Pool::get();
    const size_t w = pool.size();
    if (n < 64 || w == 1) {
        f(size_t(0), n);
        return;
    }
    const size_t per = (n + w - 1) / w;
    pool.run(w, [&](size_t t) {
        const size_t a = t * per, b = std::min(n, a + per);
        if (a < b) f(a, b);
    });
}

// f(begin, end) over [0, n) when the work (n items of `cost` flops) is worth sharing
// out, else serially. Range bounds are multiples of `align` items (8 doubles: no two
// workers write one cache line).
template <typename F>
static void parallel_if(size_t n, size_t cost, F &&f, size_t align = 1) {
    if (n * cost < (size_t(1) << 15) || n < 2) {
        f(size_t(0), n);
        return;
    }
    Pool &pool = Pool::get();
    size_t per = (n + pool.size() - 1) / pool.size();
    per = (per + align - 1) / align * align;
    const size_t w = (n + per - 1) / per;
    pool.run(w, [&](size_t t) {
        const size_t a = t * per, b = std::min(n, a + per);
        if (a < b) f(a, b);
    });
}

// sqrt(a^2 + b^2): directly when neither square can overflow or lose all its bits,
// else std::hypot.
static inline double safe_hypot(double a, double b) {
    const double m = std::max(std::fabs(a), std::fabs(b));
    if (m < 1e150 && m > 1e-150) return std::sqrt(a * a + b * b);
    return std::hypot(a, b);
}

// Dot product in four interleaved partial sums (a fixed order: the same bits on every
// run), which keeps four additions in flight.
static inline double dot4(const double *a, const double *b, size_t n) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    size_t i = 0;
    for (; i + 4 <= n; i += 4) {
        s0 += a[i] * b[i];
        s1 += a[i + 1] * b[i + 1];
        s2 += a[i + 2] * b[i + 2];
        s3 += a[i + 3] * b[i + 3];
    }
    for (; i < n; i++) s0 += a[i] * b[i];
    return (s0 + s2) + (s1 + s3);
}

Mat Mat::beside(const Mat &right) const {
    Mat out(rows, cols + right.cols);
    for (size_t r = 0; r < rows; r++) {
        std::copy(row(r), row(r) + cols, out.row(r));
        std::copy(right.row(r), right.row(r) + right.cols, out.row(r) + cols);
    }
    return out;
}

// out (p x q) = A^T B, A m x p, B m x q.
Mat mul_tn(const Mat &a, const Mat &b) {
    Mat out(a.cols, b.cols);
    parallel_for(a.cols, [&](size_t i0, size_t i1) {
        for (size_t r = 0; r < a.rows; r++) {
            const double *ar = a.row(r);
            const double *br = b.row(r);
            for (size_t i = i0; i < i1; i++) {
                double f = ar[i];
                double *o = out.row(i);
                for (size_t j = 0; j < b.cols; j++) o[j] += f * br[j];
            }
        }
    });
    return out;
}

// out (m x q) = A B, A m x p, B p x q.
Mat mul_nn(const Mat &a, const Mat &b) {
    Mat out(a.rows, b.cols);
    parallel_for(a.rows, [&](size_t r0, size_t r1) {
        for (size_t r = r0; r < r1; r++) {
            const double *ar = a.row(r);
            double *o = out.row(r);
            for (size_t k = 0; k < a.cols; k++) {
                double f = ar[k];
                const double *br = b.row(k);
                for (size_t j = 0; j < b.cols; j++) o[j] += f * br[j];
            }
        }
    });
    return out;
}

static double frobenius(const Mat &m) {
    // Scaled sum of squares: no overflow for large entries.
    double scale = 0.0, ssq = 1.0;
    for (double v : m.v) {
        if (v != 0.0) {
            double a = std::fabs(v);
            if (scale < a) {
                ssq = 1.0 + ssq * (scale / a) * (scale / a);
                scale = a;
            } else {
                ssq += (a / scale) * (a / scale);
            }
        }
    }
    return scale * std::sqrt(ssq);
}

// Thin Householder QR of a (m x n, m >= n): returns Q (m x n, orthonormal columns)
// and R (n x n, upper triangular) with a = Q R. Works on the columns stored
// contiguously (the transpose), so every reflection streams memory.
static void householder_qr(const Mat &a, Mat &q, Mat &r) {
    const size_t m = a.rows, n = a.cols;
    std::vector<double> col(n * m);  // col[c * m + i] = a(i, c)
    for (size_t i = 0; i < m; i++)
        for (size_t c = 0; c < n; c++) col[c * m + i] = a.at(i, c);
    std::vector<double> betas(n, 0.0);
    std::vector<double> vs(n * m, 0.0);  // reflection vectors, v_j in vs[j * m + j ..]
    r = Mat(n, n);
    for (size_t j = 0; j < n; j++) {
        double *cj = &col[j * m];
        const double norm = std::sqrt(dot4(cj + j, cj + j, m - j));
        double *v = &vs[j * m];
        for (size_t i = j; i < m; i++) v[i] = cj[i];
        if (norm == 0.0) continue;
        const double alpha = v[j] > 0 ? -norm : norm;
        v[j] -= alpha;
        const double vtv = dot4(v + j, v + j, m - j);
        const double beta = vtv > 0 ? 2.0 / vtv : 0.0;
        betas[j] = beta;
        // Apply (I - beta v v^T) to columns j..n-1 (rows j..m-1).
        parallel_if(n - j, 4 * (m - j), [&](size_t c0, size_t c1) {
            for (size_t c = j + c0; c < j + c1; c++) {
                double *cc = &col[c * m];
                const double sdot = beta * dot4(v + j, cc + j, m - j);
                for (size_t i = j; i < m; i++) cc[i] -= sdot * v[i];
            }
        });
    }
    for (size_t i = 0; i < n; i++)
        for (size_t c = i; c < n; c++) r.at(i, c) = col[c * m + i];
    // Q = H_0 H_1 ... H_{n-1} [I; 0], applied from the last reflection back, on
    // contiguous columns.
    std::vector<double> qc(n * m, 0.0);
    for (size_t i = 0; i < n; i++) qc[i * m + i] = 1.0;
    for (size_t jj = n; jj-- > 0;) {
        const double beta = betas[jj];
        if (beta == 0.0) continue;
        const double *v = &vs[jj * m];
        parallel_if(n, 4 * (m - jj), [&](size_t c0, size_t c1) {
            for (size_t c = c0; c < c1; c++) {
                double *cc = &qc[c * m];
                const double sdot = beta * dot4(v + jj, cc + jj, m - jj);
                for (size_t i = jj; i < m; i++) cc[i] -= sdot * v[i];
            }
        });
    }
    q = Mat(m, n);
    for (size_t i = 0; i < m; i++)
        for (size_t c = 0; c < n; c++) q.at(i, c) = qc[c * m + i];
}

// One-sided Jacobi SVD of a square a (n x n): a = U diag(s) V^T with s descending.
// Columns of U whose singular value is 0 are left as zeros.
static void jacobi_svd(const Mat &a, Mat &u, std::vector<double> &s) {
    size_t n = a.rows;
    // Work on columns: store the transpose so each column is a contiguous row.
    std::vector<std::vector<double>> col(n, std::vector<double>(n));
    for (size_t i = 0; i < n; i++)
        for (size_t j = 0; j < n; j++) col[j][i] = a.at(i, j);
    // Rotate while a pair's cosine exceeds n eps (dgesvj's sqrt(m) eps criterion, with
    // n for m): below it the columns are orthogonal to working precision.
    const double tol = (double)n * 2.220446049250313e-16;
    for (int sweep = 0; sweep < 80; sweep++) {
        bool rotated = false;
        for (size_t p = 0; p + 1 < n; p++) {
            for (size_t qq = p + 1; qq < n; qq++) {
                const std::vector<double> &cp = col[p], &cq = col[qq];
                const double alpha = dot4(cp.data(), cp.data(), n), beta = dot4(cq.data(), cq.data(), n),
                             gamma = dot4(cp.data(), cq.data(), n);
                if (gamma == 0.0 || std::fabs(gamma) <= tol * std::sqrt(alpha * beta)) continue;
                rotated = true;
                double zeta = (beta - alpha) / (2.0 * gamma);
                double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
                std::vector<double> &xp = col[p], &xq = col[qq];
                for (size_t i = 0; i < n; i++) {
                    double x = xp[i], y = xq[i];
                    xp[i] = c * x - sn * y;
                    xq[i] = sn * x + c * y;
                }
            }
        }
        if (!rotated) break;
    }
    std::vector<double> norms(n);
    for (size_t j = 0; j < n; j++) {
        double ss = 0;
        for (double x : col[j]) ss += x * x;
        norms[j] = std::sqrt(ss);
    }
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return norms[x] > norms[y]; });
    u = Mat(n, n);
    s.assign(n, 0.0);
    for (size_t k = 0; k < n; k++) {
        size_t j = order[k];
        s[k] = norms[j];
        if (norms[j] > 0)
            for (size_t i = 0; i < n; i++) u.at(i, k) = col[j][i] / norms[j];
    }
}

// A lower bound of the smallest singular value of an upper triangular r (n x n):
// 1 / |r^-1|_F, with r^-1 by back substitution; 0 when a diagonal entry is 0 or the
// inverse is not finite.
static double min_singular_bound(const Mat &r) {
    const size_t n = r.rows;
    if (n == 0) return INFINITY;
    std::vector<double> inv(n * n, 0.0);  // column c of r^-1 at inv[c * n ..]
    for (size_t c = 0; c < n; c++) {
        double *x = &inv[c * n];
        for (size_t i = c + 1; i-- > 0;) {
            const double diag = r.at(i, i);
            if (diag == 0.0) return 0.0;
            double acc = i == c ? 1.0 : 0.0;
            for (size_t j = i + 1; j <= c; j++) acc -= r.at(i, j) * x[j];
            x[i] = acc / diag;
        }
    }
    double ss = 0.0;
    for (double v : inv) ss += v * v;
    if (!std::isfinite(ss) || ss == 0.0) return 0.0;
    return 1.0 / std::sqrt(ss);
}

// Gram matrix a^T a (n x n) of a tall a (m x n): partial Grams of 64-row chunks in
// parallel, added in chunk order (the same bits on any thread count).
static Mat gram(const Mat &a) {
    const size_t m = a.rows, n = a.cols, CH = 64, nch = (m + CH - 1) / CH;
    std::vector<double> part(nch * n * n, 0.0);
    parallel_if(nch, CH * n * n, [&](size_t c0, size_t c1) {
        for (size_t c = c0; c < c1; c++) {
            double *g = &part[c * n * n];
            for (size_t r = c * CH; r < std::min(m, (c + 1) * CH); r++) {
                const double *x = a.row(r);
                for (size_t i = 0; i < n; i++) {
                    const double xi = x[i];
                    double *gi = g + i * n;
                    for (size_t j = i; j < n; j++) gi[j] += xi * x[j];
                }
            }
        }
    });
    Mat g(n, n);
    for (size_t c = 0; c < nch; c++)
        for (size_t i = 0; i < n; i++)
            for (size_t j = i; j < n; j++) g.at(i, j) += part[c * n * n + i * n + j];
    for (size_t i = 0; i < n; i++)
        for (size_t j = 0; j < i; j++) g.at(i, j) = g.at(j, i);
    return g;
}

// g = r^T r with r upper triangular; false when a pivot is not positive.
bool cholesky_factor(const Mat &g, Mat &r) {
    const size_t n = g.rows;
    r = Mat(n, n);
    for (size_t j = 0; j < n; j++) {
        double d = g.at(j, j);
        for (size_t k = 0; k < j; k++) d -= r.at(k, j) * r.at(k, j);
        if (!(d > 0.0) || !std::isfinite(d)) return false;
        const double rjj = std::sqrt(d);
        r.at(j, j) = rjj;
        for (size_t i = j + 1; i < n; i++) {
            double v = g.at(j, i);
            for (size_t k = 0; k < j; k++) v -= r.at(k, j) * r.at(k, i);
            r.at(j, i) = v / rjj;
        }
    }
    return true;
}

// a <- a r^-1 (r upper triangular), row by row.
static void solve_right_upper(Mat &a, const Mat &r) {
    const size_t n = r.rows;
    parallel_if(a.rows, n * n, [&](size_t r0, size_t r1) {
        for (size_t i = r0; i < r1; i++) {
            double *x = a.row(i);
            for (size_t j = 0; j < n; j++) {
                double v = x[j];
                for (size_t k = 0; k < j; k++) v -= x[k] * r.at(k, j);
                x[j] = v / r.at(j, j);
            }
        }
    });
}

// |r|_F |r^-1|_F, an upper bound of the 2-norm condition number of r.
double condition_bound_of(const Mat &r) {
    const double lo = min_singular_bound(r);
    return lo > 0.0 ? frobenius(r) / lo : INFINITY;
}

// Thin QR by Cholesky QR twice: a = q r when a is well conditioned (the first
// factor's condition bound at most 1e6, where two passes leave q orthonormal to
// working precision); false otherwise. Four parallel sweeps over a in place of a
// Householder QR's two per column.
static bool cholesky_qr2(const Mat &a, Mat &q, Mat &r) {
    Mat r1, r2;
    if (!cholesky_factor(gram(a), r1) || condition_bound_of(r1) > 1e6) return false;
    q = a;
    solve_right_upper(q, r1);
    if (!cholesky_factor(gram(q), r2)) return false;
    solve_right_upper(q, r2);
    const size_t n = r1.rows;
    r = Mat(n, n);
    for (size_t i = 0; i < n; i++)
        for (size_t j = i; j < n; j++) {
            double v = 0.0;
            for (size_t k = i; k <= j; k++) v += r2.at(i, k) * r1.at(k, j);
            r.at(i, j) = v;
        }
    return true;
}

// The block's scale (its Frobenius norm), 0 for a block that adds nothing.
double block_scale_of(const Mat &block) {
    const double scale = block.cols == 0 ? 0.0 : frobenius(block);
    if (!std::isfinite(scale))
        throw Error(KY_E215, "pca: orthonormalising a Krylov block did not converge (a value that is not finite)");
    return scale;
}

// The orthonormal new directions of `rest` (the block with the basis projected out).
static Mat new_directions(const Mat &rest, double scale);
static Mat keep_directions(const Mat &q, const Mat &r, double scale);

Mat extend_basis(const Mat &basis, const Mat &block, DeviceBasis *dev) {
    if (dev) {
        KY_CHECK(dev->cols == basis.cols, KY_E505, "pca: the device basis is out of step (internal)");
        return extend_basis(*dev, block);
    }
    const double scale = block_scale_of(block);
    if (scale == 0.0) return Mat(block.rows, 0);
    Mat rest = block;
    if (basis.cols > 0) {
        for (int pass = 0; pass < 2; pass++) {
            Mat proj = mul_tn(basis, rest);   // bc x bw
            Mat back = mul_nn(basis, proj);   // rows x bw
            for (size_t i = 0; i < rest.v.size(); i++) rest.v[i] -= back.v[i];
        }
    }
    return new_directions(rest, scale);
}

Mat extend_basis(DeviceBasis &dev, const Mat &block) {
    const double scale = block_scale_of(block);
    if (scale == 0.0) return Mat(block.rows, 0);
    KY_CHECK(block.rows >= block.cols, KY_E505, "pca: a Krylov block wider than the genes (internal)");
    if (block.cols > 64) {
        Mat rest = block;
        if (dev.cols > 0) dev.project_out(rest);
        return new_directions(rest, scale);
    }
    // The projection and Cholesky QR twice on the device (the block stays there), the
    // small factors on the host; a block the first factor finds ill conditioned goes
    // to the host's Householder QR.
    dev.project_out_resident(block);
    Mat q, r;
    if (!dev.cholesky_qr2(block.cols, q, r)) {
        Mat rest = q;  // the projected block
        return new_directions(rest, scale);
    }
    return keep_directions(q, r, scale);
}

size_t extend_basis_resident(DeviceBasis &dev, size_t width, double scale) {
    if (!std::isfinite(scale))
        throw Error(KY_E215, "pca: orthonormalising a Krylov block did not converge (a value that is not finite)");
    if (scale == 0.0 || width == 0) return 0;
    KY_CHECK(dev.rows >= width, KY_E505, "pca: a Krylov block wider than the genes (internal)");
    if (width > 64) {
        Mat rest = dev.get_block(width);
        if (dev.cols > 0) dev.project_out(rest);
        const Mat q = new_directions(rest, scale);
        dev.put_block(q);
        return q.cols;
    }
    dev.project_resident(width);
    Mat q, r;
    if (!dev.cholesky_qr2(width, q, r, false)) {
        const Mat out = new_directions(q, scale);  // q: the projected block
        dev.put_block(out);
        return out.cols;
    }
    // Full rank keeps Q itself, where it is (keep_directions' first test).
    if (min_singular_bound(r) > DROP_TOLERANCE * scale) return width;
    const Mat out = keep_directions(dev.get_block(width), r, scale);
    dev.put_block(out);
    return out.cols;
}

static Mat new_directions(const Mat &rest, double scale) {
    const size_t rows = rest.rows;
    size_t n = rest.cols;
    Mat q, r;
    KY_CHECK(rows >= n, KY_E505, "pca: a Krylov block wider than the genes (internal)");
    // A well-conditioned block (every Krylov block of a PCA here) by Cholesky QR twice,
    // any other by Householder.
    if (!cholesky_qr2(rest, q, r)) householder_qr(rest, q, r);
    return keep_directions(q, r, scale);
}

// The block's directions from its thin QR (q, r): q itself when r is provably of full
// rank, else q times the right singular vectors of r above the drop tolerance.
static Mat keep_directions(const Mat &q, const Mat &r, double scale) {
    const size_t rows = q.rows, n = q.cols;
    Mat ur;
    std::vector<double> s;
    // Full rank (every singular value of R above the drop tolerance) keeps the whole
    // block, and the new columns may then be Q itself: span(Q U) = span(Q) for a square
    // orthogonal U, and the passes and the Rayleigh-Ritz step depend on the basis only
    // through its span. sigma_min(R) >= 1 / |R^-1|_F (back substitution) proves it
    // without the SVD; otherwise the SVD decides which directions to drop.
    if (min_singular_bound(r) > DROP_TOLERANCE * scale) return q;
    jacobi_svd(r, ur, s);
    size_t keep = 0;
    while (keep < s.size() && s[keep] > DROP_TOLERANCE * scale) keep++;
    keep = std::min(keep, rows);
    Mat ukeep(n, keep);
    for (size_t i = 0; i < n; i++)
        for (size_t j = 0; j < keep; j++) ukeep.at(i, j) = ur.at(i, j);
    return mul_nn(q, ukeep);
}

// A barrier for the tasks of one pool run (spinning): every task of a run with no more
// tasks than the pool's active threads is running at once, so all of them arrive.
struct SpinBarrier {
    alignas(64) std::atomic<unsigned> arrived{0};
    alignas(64) std::atomic<unsigned> phase{0};
    unsigned parties;
    explicit SpinBarrier(unsigned n) : parties(n) {}
    void wait() {
        if (parties <= 1) return;
        const unsigned ph = phase.load(std::memory_order_acquire);
        if (arrived.fetch_add(1, std::memory_order_acq_rel) + 1 == parties) {
            arrived.store(0, std::memory_order_relaxed);
            phase.fetch_add(1, std::memory_order_release);
        } else {
            while (phase.load(std::memory_order_acquire) == ph) __builtin_ia32_pause();
        }
    }
};

// Householder reduction of the symmetric a (n x n, full row-major storage, updated in
// place) to tridiagonal form: d the diagonal, e[j] the coupling of j and j + 1 (e[n -
// 1] = 0), vs[j] / betas[j] the reflection of step j (beta 0: none). Step j: v = the
// row below the diagonal with its first element moved to -sign(v0)|v| ... the usual
// I - beta u u^T, p = beta A' v, K = beta/2 v.p, w = p - K v, A' -= v w^T + w v^T. The
// update of step j and the product p of step j + 1 are one sweep over the trailing
// rows (each row updated, then dotted with the next reflection's vector), cut over a
// team of threads with one barrier a step: each thread forms the small vectors (w, the
// next v) itself from the same values, so all agree to the bit, and every element is
// computed as the one-pass-a-product form would (the same bits whatever the team).
static void tridiagonalize(Mat &A, std::vector<double> &d, std::vector<double> &e,
                           std::vector<std::vector<double>> &vs, std::vector<double> &betas) {
    const size_t n = A.rows;
    d.assign(n, 0.0);
    e.assign(n, 0.0);
    vs.assign(n, {});
    betas.assign(n, 0.0);
    if (n == 0) return;
    d[0] = A.at(0, 0);
    if (n <= 2) {
        if (n == 2) {
            d[1] = A.at(1, 1);
            e[0] = A.at(1, 0);
        }
        return;
    }
    // Team: worth it while the trailing rows are many; the last steps run on one thread.
    constexpr size_t TEAM_ROWS = 48;
    Pool &pool = Pool::get();
    const unsigned team = (unsigned)std::min<size_t>({(size_t)pool.active(), (size_t)8, n / TEAM_ROWS});
    const unsigned T = std::max(1u, team);
    SpinBarrier bar(T);
    std::vector<double> pbuf[2] = {std::vector<double>(n), std::vector<double>(n)};
    // v (m values, the row's part below the diagonal) into the reflection's vector; false
    // when the row is zero (no reflection).
    const auto form = [](const double *row, size_t m, std::vector<double> &v, double &beta, double &alpha) {
        v.assign(row, row + m);
        const double norm = std::sqrt(dot4(v.data(), v.data(), m));
        if (norm == 0.0) {
            beta = 0.0;
            alpha = 0.0;
            return false;
        }
        alpha = v[0] > 0 ? -norm : norm;
        v[0] -= alpha;
        beta = 2.0 / dot4(v.data(), v.data(), m);
        return true;
    };
    const auto task = [&](size_t t) {
        std::vector<double> v, vn, w(n), u(n);
        double beta = 0.0, alpha = 0.0, betan = 0.0, alphan = 0.0;
        unsigned live = T;  // the team's size at this step (1 once the trailing rows are few)
        const auto range = [&](size_t lo, size_t &r0, size_t &r1) {
            const size_t cnt = n - lo;
            r0 = lo + cnt * t / live;
            r1 = lo + cnt * (t + 1) / live;
        };
        bool have = form(A.row(0) + 1, n - 1, v, beta, alpha);
        if (t == 0) {
            vs[0] = v;
            betas[0] = beta;
            e[0] = have ? alpha : A.at(0, 1);
        }
        if (have) {
            size_t r0, r1;
            range(1, r0, r1);
            for (size_t r = r0; r < r1; r++) pbuf[0][r - 1] = beta * dot4(A.row(r) + 1, v.data(), n - 1);
        }
        bar.wait();
        for (size_t j = 0; j + 2 < n; j++) {
            const size_t m = n - j - 1;  // trailing size of step j
            if (live > 1 && m < TEAM_ROWS) {
                // The rest on thread 0 (the barrier above saw every thread's last rows).
                if (t != 0) return;
                live = 1;
            }
            const double *pj = pbuf[j % 2].data();
            double *pn = pbuf[(j + 1) % 2].data();
            if (have) {
                const double K = 0.5 * beta * dot4(v.data(), pj, m);
                for (size_t i = 0; i < m; i++) w[i] = pj[i] - K * v[i];
            }
            // Row j + 1 (columns j + 1 ..) after this step's update.
            const double *src = A.row(j + 1) + j + 1;
            for (size_t c = 0; c < m; c++) u[c] = have ? src[c] - (v[0] * w[c] + w[0] * v[c]) : src[c];
            const bool step_next = j + 3 < n;
            const bool next = step_next && form(u.data() + 1, m - 1, vn, betan, alphan);
            if (t == 0) {
                d[j + 1] = u[0];
                if (step_next) {
                    vs[j + 1] = vn;
                    betas[j + 1] = betan;
                    e[j + 1] = next ? alphan : u[1];
                } else {
                    e[j + 1] = u[1];
                }
            }
            // Rows j + 2 .. : columns j + 2 .. updated, then dotted with the next vector
            // (four interleaved partial sums, as dot4).
            size_t r0, r1;
            range(j + 2, r0, r1);
            for (size_t r = r0; r < r1; r++) {
                double *ar = A.row(r) + j + 2;
                const size_t mm = m - 1;
                if (have) {
                    const double vr = v[r - j - 1], wr = w[r - j - 1];
                    const double *wc = w.data() + 1, *vc = v.data() + 1;
                    for (size_t c = 0; c < mm; c++) ar[c] -= vr * wc[c] + wr * vc[c];
                }
                if (next) pn[r - j - 2] = betan * dot4(ar, vn.data(), mm);
            }
            if (live > 1) bar.wait();
            std::swap(v, vn);
            beta = betan;
            alpha = alphan;
            have = next;
        }
        if (t == 0) d[n - 1] = A.at(n - 1, n - 1);
    };
    if (T == 1)
        task(0);
    else
        pool.run(T, task);
    e[n - 1] = 0.0;
}

// Symmetric eigendecomposition a (n x n) = V diag(d) V^T: every eigenvalue (d
// ascending) and the eigenvectors of the `want` largest (vout, n x want, largest
// first): the Householder tridiagonalisation above, then the implicit QL method; only
// the wanted eigenvectors are formed, from the recorded rotations and reflections.
static void symmetric_eigen(const Mat &a, size_t want, std::vector<double> &d, Mat &vout) {
    const size_t n = a.rows;
    d.assign(n, 0.0);
    std::vector<double> e(n, 0.0);
    if (n == 0) { vout = Mat(0, 0); return; }
    Mat A = a;
    std::vector<std::vector<double>> vs;
    std::vector<double> betas;
    tridiagonalize(A, d, e, vs, betas);
    // Implicit QL on the tridiagonal matrix (e[i] couples i and i + 1), its rotations
    // recorded: the eigenvector matrix is Q R_1 ... R_N (Q the reflections above, R_t
    // the t-th rotation acting on columns i, i + 1).
    struct Rotation {
        size_t i;
        double c, s;
    };
    std::vector<Rotation> rotations;
    double f = 0.0, tst1 = 0.0;
    const double eps = 2.220446049250313e-16;
    for (size_t l = 0; l < n; l++) {
        tst1 = std::max(tst1, std::fabs(d[l]) + std::fabs(e[l]));
        size_t m = l;
        while (m < n) {
            if (std::fabs(e[m]) <= eps * tst1) break;
            m++;
        }
        if (m > l) {
            int iter = 0;
            do {
                iter++;
                if (iter > 200)
                    throw Error(KY_E215, "pca: the Rayleigh-Ritz eigendecomposition did not converge");
                double g = d[l];
                double p = (d[l + 1] - g) / (2.0 * e[l]);
                double r = safe_hypot(p, 1.0);
                if (p < 0) r = -r;
                d[l] = e[l] / (p + r);
                d[l + 1] = e[l] * (p + r);
                double dl1 = d[l + 1];
                double h = g - d[l];
                for (size_t i = l + 2; i < n; i++) d[i] -= h;
                f += h;
                p = d[m];
                double c = 1.0, c2 = c, c3 = c;
                double el1 = e[l + 1];
                double s = 0.0, s2 = 0.0;
                for (size_t ii = m; ii-- > l;) {
                    c3 = c2;
                    c2 = c;
                    s2 = s;
                    g = c * e[ii];
                    h = c * p;
                    r = safe_hypot(p, e[ii]);
                    e[ii + 1] = s * r;
                    s = e[ii] / r;
                    c = p / r;
                    p = c * d[ii] - s * g;
                    d[ii + 1] = h + s * (c * g + s * d[ii]);
                    rotations.push_back({ii, c, s});
                }
                p = -s * s2 * c3 * el1 * e[l] / dl1;
                e[l] = s * p;
                d[l] = c * p;
            } while (std::fabs(e[l]) > eps * tst1);
        }
        d[l] = d[l] + f;
        e[l] = 0.0;
    }
    // The eigenvalues ascending (ties in index order), then the `want` largest's
    // vectors, largest first: column j is Q R_1 ... R_N e_{s_j}, the rotations applied
    // to the unit vectors from the last back (each a pair of rows of the n x want
    // block), then the reflections from the first back. Every column is its own
    // computation: the block is cut into one slice a thread (whole 64-byte lines of each
    // row: a multiple of 8 columns), the slices run on the pool.
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return d[x] < d[y]; });
    std::vector<double> ds(n);
    for (size_t k = 0; k < n; k++) ds[k] = d[order[k]];
    want = std::min(want, n);
    const size_t lines = (want + 7) / 8;  // 8-column lines a row
    const size_t threads = n >= 64 ? std::min<size_t>(Pool::get().active(), lines) : 1;
    const size_t sw = (lines + threads - 1) / threads * 8;  // slice width
    const size_t slices = (lines * 8 + sw - 1) / sw;
    const size_t ys = slices * sw;  // row stride
    std::vector<double> store(n * ys + 8, 0.0);
    double *Y = store.data();
    while (reinterpret_cast<uintptr_t>(Y) % 64 != 0) Y++;  // rows on 64-byte lines
    for (size_t j = 0; j < want; j++) Y[order[n - 1 - j] * ys + j] = 1.0;
    const auto slice = [&](size_t sl) {
        const size_t c0 = sw * sl;
        for (size_t t = rotations.size(); t-- > 0;) {
            const Rotation &rot = rotations[t];
            double *__restrict y0 = Y + rot.i * ys + c0;
            double *__restrict y1 = Y + (rot.i + 1) * ys + c0;
            const double c = rot.c, sn = rot.s;
            for (size_t j = 0; j < sw; j++) {
                const double a0 = y0[j], a1 = y1[j];
                y0[j] = c * a0 + sn * a1;
                y1[j] = c * a1 - sn * a0;
            }
        }
        std::vector<double> acc(sw);
        for (size_t jj = n >= 2 ? n - 2 : 0; jj-- > 0;) {
            if (betas[jj] == 0.0) continue;
            const std::vector<double> &v = vs[jj];
            const size_t m = n - jj - 1;
            std::fill(acc.begin(), acc.end(), 0.0);
            for (size_t r = 0; r < m; r++) {
                const double vr = v[r];
                const double *y = Y + (jj + 1 + r) * ys + c0;
                for (size_t j = 0; j < sw; j++) acc[j] += vr * y[j];
            }
            for (size_t j = 0; j < sw; j++) acc[j] *= betas[jj];
            for (size_t r = 0; r < m; r++) {
                const double vr = v[r];
                double *y = Y + (jj + 1 + r) * ys + c0;
                for (size_t j = 0; j < sw; j++) y[j] -= acc[j] * vr;
            }
        }
    };
    if (slices > 1)
        Pool::get().run(slices, slice);
    else
        for (size_t sl = 0; sl < slices; sl++) slice(sl);
    vout = Mat(n, want);
    for (size_t r = 0; r < n; r++) std::copy(Y + r * ys, Y + r * ys + want, vout.row(r));
    d = ds;
}

// pca_dense.rs:215-252.
static Mat complement(const Mat &columns, size_t count) {
    Mat all = columns;
    Mat added(columns.rows, 0);
    for (size_t c = 0; c < count; c++) {
        size_t gene = 0;
        double least = INFINITY;
        for (size_t r = 0; r < all.rows; r++) {
            double cov = 0.0;
            for (size_t j = 0; j < all.cols; j++) cov += all.at(r, j) * all.at(r, j);
            if (cov < least) {
                least = cov;
                gene = r;
            }
        }
        std::vector<double> v(all.rows, 0.0);
        v[gene] = 1.0;
        for (int pass = 0; pass < 2; pass++) {
            std::vector<double> coef(all.cols, 0.0);
            for (size_t j = 0; j < all.cols; j++)
                for (size_t r = 0; r < all.rows; r++) coef[j] += all.at(r, j) * v[r];
            for (size_t r = 0; r < all.rows; r++) {
                double s = 0.0;
                for (size_t j = 0; j < all.cols; j++) s += all.at(r, j) * coef[j];
                v[r] -= s;
            }
        }
        double norm = 0.0;
        for (double x : v) norm += x * x;
        norm = std::sqrt(norm);
        Mat col(all.rows, 1);
        for (size_t r = 0; r < all.rows; r++) col.at(r, 0) = v[r] / norm;
        all = all.beside(col);
        added = added.beside(col);
    }
    return added;
}

// pca_dense.rs:256-273: each column's largest absolute value (the first on ties)
// made positive.
static void flip_signs(Mat &m) {
    for (size_t c = 0; c < m.cols; c++) {
        double largest = 0.0, sign = 1.0;
        for (size_t r = 0; r < m.rows; r++) {
            double v = m.at(r, c);
            if (std::fabs(v) > largest) {
                largest = std::fabs(v);
                sign = v < 0 ? -1.0 : 1.0;
            }
        }
        if (sign < 0)
            for (size_t r = 0; r < m.rows; r++) m.at(r, c) = -m.at(r, c);
    }
}

// The Rayleigh-Ritz step over a basis of `cols` columns (of `rows` genes): B^T C and
// B U from the host matrices, or from `dev` when given.
static void rayleigh_ritz_impl(size_t rows, size_t cols, const Mat *basis, const Mat *products, size_t k,
                               std::vector<double> &eigenvalues, Mat &loadings, DeviceBasis *dev) {
    loadings = Mat(rows, 0);
    eigenvalues.clear();
    if (cols > 0) {
        Mat t = dev ? dev->rayleigh_matrix() : mul_tn(*basis, *products);
        size_t p = t.rows;
        Mat sym(p, p);
        for (size_t i = 0; i < p; i++)
            for (size_t j = 0; j < p; j++) sym.at(i, j) = 0.5 * (t.at(i, j) + t.at(j, i));
        for (double v : sym.v)
            if (!std::isfinite(v))
                throw Error(KY_E215, "pca: the Rayleigh-Ritz eigendecomposition did not converge (a value that is not finite)");
        std::vector<double> d;
        Mat chosen;
        size_t taken = std::min(k, p);
        symmetric_eigen(sym, taken, d, chosen);
        for (size_t c = 0; c < taken; c++) eigenvalues.push_back(std::max(d[p - 1 - c], 0.0));
        loadings = dev ? dev->times(chosen) : mul_nn(*basis, chosen);
    }
    size_t missing = k - eigenvalues.size();
    if (missing > 0) {
        loadings = loadings.beside(complement(loadings, missing));
        eigenvalues.resize(k, 0.0);
    }
    flip_signs(loadings);
}

void rayleigh_ritz(const Mat &basis, const Mat &products, size_t k, std::vector<double> &eigenvalues,
                   Mat &loadings, DeviceBasis *dev) {
    if (dev) KY_CHECK(dev->cols == basis.cols, KY_E505, "pca: the device basis is out of step (internal)");
    rayleigh_ritz_impl(basis.rows, basis.cols, &basis, &products, k, eigenvalues, loadings, dev);
}

void rayleigh_ritz(DeviceBasis &dev, size_t k, std::vector<double> &eigenvalues, Mat &loadings) {
    rayleigh_ritz_impl(dev.rows, dev.cols, nullptr, nullptr, k, eigenvalues, loadings, &dev);
}

}  // namespace ky
