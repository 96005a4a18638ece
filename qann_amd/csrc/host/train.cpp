// Coarse and PQ training with the assign steps handed to the GPU (train.hpp).
//
// The sequence of random draws, the order of every float sum and the fused
// multiply-adds are the reference's as g++ -O3 -mavx2 -mfma compiles
// src/kmeans.cpp: its only contractions in these functions are the EMA step
// c + w * (mean - c) and the cosine schedule's 0.1 + 0.45 * (1 + cos), written
// below as explicit fma; this file is built without FMA contraction. The
// random draws go through the same <random> distributions and std::shuffle
// over the same types, so the same libstdc++ yields the same draws.
#include "train.hpp"

#include "parallel.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <queue>
#include <random>
#include <stdexcept>
#include <utility>

namespace qh {

float l2_avx2(const float* a, const float* b, uint64_t n) {
    // four 8-lane accumulators over steps of 32, then steps of 8 into the first
    float acc[4][8] = {};
    uint64_t i = 0;
    for (; i + 32 <= n; i += 32)
        for (int r = 0; r < 4; ++r)
            for (int l = 0; l < 8; ++l) {
                const float d = a[i + 8 * r + l] - b[i + 8 * r + l];
                acc[r][l] = std::fma(d, d, acc[r][l]);
            }
    for (; i + 8 <= n; i += 8)
        for (int l = 0; l < 8; ++l) {
            const float d = a[i + l] - b[i + l];
            acc[0][l] = std::fma(d, d, acc[0][l]);
        }
    float v[8];
    for (int l = 0; l < 8; ++l) v[l] = (acc[0][l] + acc[1][l]) + (acc[2][l] + acc[3][l]);
    // hsum256: the halves added, then two horizontal adds
    float result = ((v[0] + v[4]) + (v[1] + v[5])) + ((v[2] + v[6]) + (v[3] + v[7]));
    // the scalar tail as compiled: with four or more left, four squares
    // rounded and added in turn, then the rest fused
    const uint64_t r = n - i;
    uint64_t j = 0;
    if (r >= 4)
        for (; j < 4; ++j) {
            const float d = a[i + j] - b[i + j];
            const float sq = d * d;
            result = result + sq;
        }
    for (; j < r; ++j) {
        const float d = a[i + j] - b[i + j];
        result = std::fma(d, d, result);
    }
    return result;
}

namespace {

constexpr double kPi = 3.141592653589793; // std::numbers::pi

std::mt19937 seeded(std::optional<uint32_t> seed) { return std::mt19937(seed ? *seed : std::random_device{}()); }

// sums[c][l] += row[l] for every row, in row order per (c, l); the dimensions
// are split over threads, which leaves each sum's order unchanged.
void sum_by_cluster(const float* x, uint64_t n, uint64_t stride, uint64_t d, const uint32_t* label, float* sums) {
    constexpr uint64_t kDims = 64;
    parallel_for((d + kDims - 1) / kDims, [&](size_t c) {
        const uint64_t lo = c * kDims, hi = std::min<uint64_t>(d, lo + kDims);
        for (uint64_t r = 0; r < n; ++r) {
            float* dst = sums + static_cast<uint64_t>(label[r]) * d;
            const float* src = x + r * stride;
            for (uint64_t l = lo; l < hi; ++l) dst[l] += src[l];
        }
    });
}

// Unit-length copy of each row (a double sum of squares, one float scale);
// zero rows stay zero (src/kmeans.cpp:248-256).
void unit_rows(const float* c, uint64_t k, uint64_t d, float* out) {
    for (uint64_t j = 0; j < k; ++j) {
        double sq = 0;
        for (uint64_t l = 0; l < d; ++l) sq += double(c[j * d + l]) * c[j * d + l];
        const float scale = sq > 0 ? static_cast<float>(1.0 / std::sqrt(sq)) : 0.0f;
        for (uint64_t l = 0; l < d; ++l) out[j * d + l] = c[j * d + l] * scale;
    }
}

uint64_t nearest_power_of_two(double v) {
    if (v <= 1) return 1;
    const uint64_t low = uint64_t{1} << static_cast<int>(std::floor(std::log2(v)));
    const uint64_t high = 2 * low;
    return v / double(low) < double(high) / v ? low : high;
}

// One detail::kmeans run over rows x + i * stride of d floats, stepped from
// outside: the caller assigns, step() moves the centroids.
class KMeansRun {
  public:
    KMeansRun(const float* x, uint64_t n, uint64_t stride, uint64_t d, uint64_t k, std::optional<uint32_t> seed,
              bool spherical)
        : x_(x), n_(n), stride_(stride), d_(d), k_(k), spherical_(spherical), cents_(k * d) {
        if (n < k) throw std::invalid_argument("kmeans: need at least k points");
        std::mt19937 rng = seeded(seed);
        // k distinct rows: a partial Fisher-Yates by uniform draws over [i, n)
        std::vector<size_t> idx(n);
        std::iota(idx.begin(), idx.end(), size_t{0});
        for (uint64_t i = 0; i < k; ++i) {
            std::uniform_int_distribution<size_t> draw(i, n - 1);
            const size_t pick = draw(rng);
            std::swap(idx[i], idx[pick]);
        }
        for (uint64_t i = 0; i < k; ++i) std::copy_n(x + idx[i] * stride, d, cents_.data() + i * d);
    }

    const std::vector<float>& centroids() const { return cents_; }

    // Means of this assignment (unit length when spherical), then every empty
    // cluster reseeded (src/kmeans.cpp:199-239).
    void step(const uint32_t* assign) {
        std::vector<float> next(k_ * d_, 0.0f);
        std::vector<size_t> count(k_, 0);
        sum_by_cluster(x_, n_, stride_, d_, assign, next.data());
        for (uint64_t r = 0; r < n_; ++r) ++count[assign[r]];
        for (uint64_t j = 0; j < k_; ++j) {
            if (count[j] == 0) continue;
            float* c = next.data() + j * d_;
            const float cnt = static_cast<float>(count[j]);
            for (uint64_t l = 0; l < d_; ++l) c[l] /= cnt;
            if (spherical_) {
                double sq = 0;
                for (uint64_t l = 0; l < d_; ++l) sq += double(c[l]) * c[l];
                if (sq > 0) {
                    const float scale = static_cast<float>(1.0 / std::sqrt(sq));
                    for (uint64_t l = 0; l < d_; ++l) c[l] *= scale;
                }
            }
        }
        reseed_empty(next.data(), assign, count);
        cents_ = std::move(next);
    }

  private:
    // reseed_empty_clusters (src/kmeans.cpp:116-170): the rows farthest from
    // their centroid (distances of at least 1e-6), nearest of them first, fill
    // the empty clusters in index order; without such rows the largest cluster
    // is split.
    void reseed_empty(float* cents, const uint32_t* assign, std::vector<size_t>& count) const {
        std::vector<uint64_t> empty;
        for (uint64_t j = 0; j < k_; ++j)
            if (count[j] == 0) empty.push_back(j);
        if (empty.empty()) return;
        using Far = std::pair<float, size_t>;
        std::priority_queue<Far, std::vector<Far>, std::greater<Far>> far;
        for (uint64_t i = 0; i < n_; ++i) {
            const float dist = l2_avx2(x_ + i * stride_, cents + static_cast<uint64_t>(assign[i]) * d_, d_);
            if (dist < 1e-6f) continue;
            if (far.size() < empty.size()) {
                far.emplace(dist, i);
            } else if (dist > far.top().first) {
                far.pop();
                far.emplace(dist, i);
            }
        }
        for (uint64_t e : empty) {
            float* dst = cents + e * d_;
            if (!far.empty()) {
                std::copy_n(x_ + far.top().second * stride_, d_, dst);
                far.pop();
                continue;
            }
            const uint64_t big = static_cast<uint64_t>(std::max_element(count.begin(), count.end()) - count.begin());
            const float eps = 1.0f / 1024.0f;
            float* src = cents + big * d_;
            for (uint64_t l = 0; l < d_; ++l) {
                dst[l] = src[l] * (1.0f + eps);
                src[l] *= (1.0f - eps);
            }
            count[e] = count[big] / 2;
            count[big] -= count[e];
        }
    }

    const float* x_;
    uint64_t n_, stride_, d_, k_;
    bool spherical_;
    std::vector<float> cents_;
};

} // namespace

Clusters lloyd_kmeans(const float* x, uint64_t n, uint64_t d, uint64_t k, uint64_t iters, std::optional<uint32_t> seed,
                      bool spherical, const NearestFn& nearest) {
    KMeansRun run(x, n, d, d, k, seed, spherical);
    Clusters out;
    out.assign.resize(n);
    for (uint64_t it = 0; it < iters; ++it) {
        nearest(x, n, d, run.centroids().data(), k, out.assign.data());
        run.step(out.assign.data());
    }
    out.centroids = run.centroids();
    nearest(x, n, d, out.centroids.data(), k, out.assign.data());
    return out;
}

Clusters emeans(const float* x, uint64_t n, uint64_t d, uint64_t k, uint64_t passes, std::optional<uint32_t> seed,
                bool spherical, const NearestFn& nearest) {
    if (n < k) throw std::invalid_argument("emeans: need at least k points");
    if (passes == 0) throw std::invalid_argument("emeans: need at least one pass");
    std::mt19937 rng = seeded(seed);

    // batch size and initial step from the fitted laws (throughput law above
    // 16 dimensions, quality law at or below), as the reference sets them
    const double T = double(passes), nd = double(n), kd = double(k);
    const bool quality = d <= 16;
    uint64_t batch = nearest_power_of_two((quality ? 0.00439175 : 0.035134) * std::pow(kd, 0.134749) *
                                          std::pow(nd, 0.876473) * std::pow(T, 0.813344));
    batch = quality ? std::min<uint64_t>(batch, n / 2) : std::min<uint64_t>(batch, nearest_power_of_two(nd / 2));
    batch = std::max<uint64_t>(1, std::min<uint64_t>(batch, n));
    const double alpha0 = std::clamp(0.8244 * std::pow(kd, -0.4617) * std::pow(double(batch), 0.9003) *
                                         std::pow(nd, -0.4498) * std::pow(T, -0.1057),
                                     1e-4, 1.0);
    const uint64_t steps_per_pass = (n + batch - 1) / batch, total_steps = passes * steps_per_pass;
    const float tau = d <= 2 ? 0.0f : 0.1f;

    // k distinct seeded rows as the starting positions
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), size_t{0});
    for (uint64_t i = 0; i < k; ++i) {
        std::uniform_int_distribution<size_t> draw(i, n - 1);
        std::swap(order[i], order[draw(rng)]);
    }
    std::vector<float> pos(k * d);
    for (uint64_t j = 0; j < k; ++j) std::copy_n(x + order[j] * d, d, pos.data() + j * d);
    if (spherical) unit_rows(pos.data(), k, d, pos.data());
    std::vector<float> mass(k, 1.0f / 65536.0f);

    std::vector<float> unit(spherical ? k * d : 0), rows(batch * d), sums(k * d), last_sums(k * d, 0.0f);
    std::vector<double> last_counts(k, 0.0);
    std::vector<uint32_t> counts(k), labels(batch);
    std::uniform_real_distribution<double> unif(0.0, 1.0);

    uint64_t step = 0;
    for (uint64_t pass = 0; pass < passes; ++pass) {
        std::shuffle(order.begin(), order.end(), rng);
        const bool last = pass + 1 == passes;
        for (uint64_t start = 0; start < n; start += batch, ++step) {
            const uint64_t b = std::min(batch, n - start);
            for (uint64_t r = 0; r < b; ++r) std::copy_n(x + order[start + r] * d, d, rows.data() + r * d);
            const float* cent = pos.data();
            if (spherical) {
                unit_rows(pos.data(), k, d, unit.data());
                cent = unit.data();
            }
            nearest(rows.data(), b, d, cent, k, labels.data());

            std::fill(counts.begin(), counts.end(), 0u);
            for (uint64_t r = 0; r < b; ++r) ++counts[labels[r]];
            std::fill(sums.begin(), sums.end(), 0.0f);
            sum_by_cluster(rows.data(), b, d, d, labels.data(), sums.data());
            if (last) {
                for (uint64_t j = 0; j < k; ++j) last_counts[j] += counts[j];
                for (uint64_t i = 0; i < k * d; ++i) last_sums[i] += sums[i];
            }

            // the step size on a cosine schedule down to 0.1 alpha0
            const double frac = double(std::min(step, total_steps)) / double(std::max<uint64_t>(1, total_steps));
            const double alpha = alpha0 * std::fma(1.0 + std::cos(kPi * frac), 0.45, 0.1);
            const bool replace = alpha == 1.0;
            if (!replace)
                for (float& mj : mass) mj = static_cast<float>(mj * (1.0 - alpha));
            const float a = static_cast<float>(alpha);
            for (uint64_t j = 0; j < k; ++j) {
                if (counts[j] == 0) continue;
                const float nj = static_cast<float>(counts[j]);
                const float grown = a * nj;
                const float new_mass = (replace ? 0.0f : mass[j]) + grown;
                const float w = grown / new_mass;
                float* c = pos.data() + j * d;
                const float* s = sums.data() + j * d;
                for (uint64_t l = 0; l < d; ++l) {
                    const float mean = s[l] / nj;
                    c[l] = w == 1.0f ? mean : std::fma(w, mean - c[l], c[l]);
                }
                mass[j] = new_mass;
            }

            // every fifth step: clusters under tau of the mean mass are refilled
            // by splitting donors drawn without replacement in proportion to mass
            if (tau > 0 && (step + 1) % 5 == 0) {
                float total = 0;
                for (float mj : mass) total += mj;
                const float threshold = total * tau / static_cast<float>(k);
                std::vector<size_t> light;
                std::vector<std::pair<double, size_t>> donors;
                for (uint64_t j = 0; j < k; ++j) {
                    if (mass[j] < threshold) light.push_back(j);
                    else if (mass[j] > 0) donors.emplace_back(0.0, j);
                }
                const size_t nsplit = std::min(light.size(), donors.size());
                if (nsplit > 0) {
                    // Gumbel keys: log mass - log(-log u)
                    for (auto& [key, j] : donors)
                        key = std::log(double(mass[j])) - std::log(-std::log(std::max(unif(rng), 1e-300)));
                    std::partial_sort(donors.begin(), donors.begin() + nsplit, donors.end(), std::greater<>());
                    for (size_t t = 0; t < nsplit; ++t) {
                        const size_t src = donors[t].second, dst = light[t];
                        float* from = pos.data() + src * d;
                        float* to = pos.data() + dst * d;
                        bool same = true;
                        for (uint64_t l = 0; l < d; ++l) {
                            const float off = (l % 2 == 0 ? 1e-4f : -1e-4f);
                            const float v = from[l];
                            from[l] = v * (1.0f - off);
                            to[l] = v * (1.0f + off);
                            same = same && from[l] == to[l];
                        }
                        if (same) to[0] = std::nextafter(to[0], std::numeric_limits<float>::infinity());
                        mass[src] = mass[dst] = mass[src] / 2.0f;
                    }
                }
            }
        }
    }

    // the last pass's means; clusters it never saw keep their position
    for (uint64_t j = 0; j < k; ++j) {
        if (last_counts[j] <= 0) continue;
        const float cnt = static_cast<float>(last_counts[j]);
        for (uint64_t l = 0; l < d; ++l) pos[j * d + l] = last_sums[j * d + l] / cnt;
    }
    if (spherical) unit_rows(pos.data(), k, d, pos.data());

    Clusters out;
    out.assign.resize(n);
    nearest(x, n, d, pos.data(), k, out.assign.data());
    out.centroids = std::move(pos);
    return out;
}

std::vector<float> pq_codebooks(const float* residuals, uint64_t n, uint64_t dim, uint64_t m, uint64_t K,
                                uint64_t iters, std::optional<uint32_t> seed, const SubspaceNearestFn& nearest) {
    if (n < K) throw std::invalid_argument("training set must have at least centroids_per_subspace_ vectors");
    // at most 256 rows per centroid, drawn by a seeded partial shuffle
    const uint64_t cap = 256 * K;
    const float* x = residuals;
    std::vector<float> sample;
    if (n > cap) {
        std::mt19937 rng = seeded(seed);
        std::vector<size_t> rows(n);
        std::iota(rows.begin(), rows.end(), size_t{0});
        sample.resize(cap * dim);
        for (uint64_t i = 0; i < cap; ++i) {
            std::uniform_int_distribution<size_t> draw(i, n - 1);
            std::swap(rows[i], rows[draw(rng)]);
            std::copy_n(residuals + rows[i] * dim, dim, sample.data() + i * dim);
        }
        x = sample.data();
        n = cap;
    }
    // one k-means run per subspace (seed + s), all stepped together so that
    // each iteration's assignments of every subspace are one device pass
    const uint64_t sub = dim / m;
    std::vector<KMeansRun> runs;
    runs.reserve(m);
    for (uint64_t s = 0; s < m; ++s)
        runs.emplace_back(x + s * sub, n, dim, sub, K,
                          seed ? std::optional<uint32_t>(*seed + static_cast<uint32_t>(s)) : std::nullopt, false);
    std::vector<float> cents(m * K * sub);
    auto gather = [&] {
        for (uint64_t s = 0; s < m; ++s) std::copy(runs[s].centroids().begin(), runs[s].centroids().end(), cents.begin() + s * K * sub);
    };
    std::vector<uint8_t> codes(n * m);
    for (uint64_t it = 0; it < iters; ++it) {
        gather();
        nearest(x, n, dim, m, cents.data(), K, codes.data());
        parallel_for(m, [&](size_t s) {
            std::vector<uint32_t> a(n);
            for (uint64_t i = 0; i < n; ++i) a[i] = codes[i * m + s];
            runs[s].step(a.data());
        });
    }
    gather();
    return cents;
}

// ---------------------------------------------------------------------------
// The PCA-balanced rotation. The reference's build contracts its double
// arithmetic as follows (read from the x86 code of the reference compiled
// here): the covariance accumulates fma(x_p, x_q, acc); the Jacobi step takes
// t from sqrt(fma(theta, theta, 1)) and c from sqrt(fma(t, t, 1)), and updates
// each pair as (fma(c, u, -(s v)), fma(s, u, c v)); the squared sums of the
// convergence test round each product of a 4-wide group and add them in
// order, fusing only the remainder. These functions are compiled with FMA
// instructions and without any other contraction.
// ---------------------------------------------------------------------------

#pragma GCC push_options
#pragma GCC optimize("fp-contract=off")

namespace {

// Sum of squares of v[0..len) onto acc, in the vectorized loop's order.
__attribute__((target("avx2,fma"))) double sum_squares(const double* v, uint64_t len, double acc) {
    const uint64_t body = len / 4 * 4;
    for (uint64_t i = 0; i < body; ++i) {
        const double sq = v[i] * v[i];
        acc = acc + sq;
    }
    for (uint64_t i = body; i < len; ++i) acc = std::fma(v[i], v[i], acc);
    return acc;
}

// symmetric_eigen (src/rotation.cpp:20-69): cyclic Jacobi sweeps over the
// pairs p < q until the off-diagonal squares fall to 1e-24 of the total.
__attribute__((target("avx2,fma"))) void jacobi_eigen(std::vector<double>& a, uint64_t d, std::vector<double>& values,
                                                      std::vector<double>& vecs) {
    vecs.assign(d * d, 0.0);
    for (uint64_t i = 0; i < d; ++i) vecs[i * d + i] = 1.0;
    const double stop = 1e-24 * sum_squares(a.data(), d * d, 0.0);
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (uint64_t p = 0; p < d; ++p) off = sum_squares(a.data() + p * d + p + 1, d - p - 1, off);
        if (off <= stop) break;
        for (uint64_t p = 0; p < d; ++p)
            for (uint64_t q = p + 1; q < d; ++q) {
                const double apq = a[p * d + q];
                if (apq == 0.0) continue;
                const double theta = (a[q * d + q] - a[p * d + p]) / (2.0 * apq);
                double t;
                if (std::abs(theta) > 1e150) {
                    t = 1.0 / (2.0 * theta);
                } else {
                    const double root = std::sqrt(std::fma(theta, theta, 1.0));
                    t = (theta >= 0 ? 1.0 : -1.0) / (std::abs(theta) + root);
                }
                const double c = 1.0 / std::sqrt(std::fma(t, t, 1.0));
                const double s = t * c;
                // columns p and q, then rows p and q, then the eigenvector columns
                auto turn = [&](double& u, double& v) {
                    const double su = s * v, cv = c * v;
                    const double nu = std::fma(c, u, -su), nv = std::fma(s, u, cv);
                    u = nu;
                    v = nv;
                };
                for (uint64_t k = 0; k < d; ++k) turn(a[k * d + p], a[k * d + q]);
                for (uint64_t k = 0; k < d; ++k) turn(a[p * d + k], a[q * d + k]);
                for (uint64_t k = 0; k < d; ++k) turn(vecs[k * d + p], vecs[k * d + q]);
            }
    }
    values.resize(d);
    for (uint64_t i = 0; i < d; ++i) values[i] = a[i * d + i];
}

// The upper triangle of sum_i (x_i - mean)^T (x_i - mean), row i after row i - 1
// for every entry, in blocks of 4,096 centred rows.
__attribute__((target("avx2,fma"))) void covariance_upper(const float* x, uint64_t n, uint64_t d,
                                                          const std::vector<double>& mean, std::vector<double>& cov) {
    constexpr uint64_t kRows = 4096;
    std::vector<double> blk(std::min(n, kRows) * d);
    for (uint64_t start = 0; start < n; start += kRows) {
        const uint64_t rows = std::min(kRows, n - start);
        for (uint64_t i = 0; i < rows; ++i)
            for (uint64_t j = 0; j < d; ++j) blk[i * d + j] = static_cast<double>(x[(start + i) * d + j]) - mean[j];
        parallel_for(d, [&](size_t p) {
            double* row = cov.data() + p * d;
            for (uint64_t i = 0; i < rows; ++i) {
                const double* b = blk.data() + i * d;
                const double bp = b[p];
                for (uint64_t q = p; q < d; ++q) row[q] = std::fma(bp, b[q], row[q]);
            }
        });
    }
}

} // namespace

std::vector<float> balanced_pca_rotation(const float* x, uint64_t n, uint64_t d, uint64_t m) {
    if (m == 0 || d % m != 0) throw std::invalid_argument("dim must be divisible by num_subspaces");
    if (n == 0) throw std::invalid_argument("training set must not be empty");
    std::vector<double> mean(d, 0.0);
    for (uint64_t i = 0; i < n; ++i)
        for (uint64_t j = 0; j < d; ++j) mean[j] += static_cast<double>(x[i * d + j]);
    for (double& v : mean) v /= static_cast<double>(n);
    std::vector<double> cov(d * d, 0.0);
    covariance_upper(x, n, d, mean, cov);
    const double scale = n > 1 ? 1.0 / static_cast<double>(n - 1) : 1.0;
    for (uint64_t p = 0; p < d; ++p)
        for (uint64_t q = p; q < d; ++q) {
            cov[p * d + q] *= scale;
            cov[q * d + p] = cov[p * d + q];
        }
    std::vector<double> values, vecs;
    jacobi_eigen(cov, d, values, vecs);

    // largest eigenvalue first (ties in index order), each to the subspace
    // with room whose sum of log eigenvalues is smallest (the first such)
    std::vector<uint64_t> order(d);
    std::iota(order.begin(), order.end(), uint64_t{0});
    std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return values[a] > values[b]; });
    const uint64_t width = d / m;
    std::vector<uint64_t> filled(m, 0);
    std::vector<double> logs(m, 0.0);
    std::vector<float> R(d * d);
    for (uint64_t e : order) {
        uint64_t best = m;
        for (uint64_t b = 0; b < m; ++b)
            if (filled[b] < width && (best == m || logs[b] < logs[best])) best = b;
        logs[best] += std::log(std::max(values[e], 1e-12));
        const uint64_t col = best * width + filled[best]++;
        for (uint64_t row = 0; row < d; ++row) R[row * d + col] = static_cast<float>(vecs[row * d + e]);
    }
    return R;
}

#pragma GCC pop_options

} // namespace qh
