// Training of an IVF index's coarse centroids and PQ codebooks, step for step
// as the reference trains them (detail::emeans and detail::kmeans,
// src/kmeans.cpp:172-414; PQCodebook::train, src/pq.cpp:41-73), so that the
// same sample and seed give the reference's centroids bit for bit. Every
// assign step -- the whole of the arithmetic that grows with n x k -- is a
// callback, which the index objects run on the GPU; the updates between them
// (sums in row order, means, EMA steps, reseeds) stay on the host.
#pragma once

#include <cstdint>
#include <functional>
#include <optional>
#include <vector>

namespace qh {

// How IVFIndex.train learns the coarse centroids (vecengine::Clustering,
// include/vecengine/ivf_index.hpp:27-30).
enum class ClusteringKind : uint8_t { EMeans = 0, Lloyd = 1 };

// The nearest of k centroids (k x d, contiguous) for each of n contiguous rows
// (n x d): the smallest exact squared L2 distance in l2_distance's order, the
// lowest index on ties.
using NearestFn =
    std::function<void(const float* rows, uint64_t n, uint64_t d, const float* cents, uint64_t k, uint32_t* out)>;

// For each of n rows (dim floats) and each of m subspaces of dim / m floats,
// the nearest of the K centroids of that subspace (cents [s][c][dim / m]),
// by the same rule; out is n x m, row-major.
using SubspaceNearestFn = std::function<void(const float* rows, uint64_t n, uint64_t dim, uint64_t m,
                                             const float* cents, uint64_t K, uint8_t* out)>;

struct Clusters {
    std::vector<float> centroids; // k x d
    std::vector<uint32_t> assign; // nearest centroid of each row under the final centroids
};

// detail::kmeans (src/kmeans.cpp:172-243): Lloyd's k-means for iters
// iterations from k distinct seeded rows; spherical keeps unit centroids.
Clusters lloyd_kmeans(const float* x, uint64_t n, uint64_t d, uint64_t k, uint64_t iters, std::optional<uint32_t> seed,
                      bool spherical, const NearestFn& nearest);

// detail::emeans (src/kmeans.cpp:265-414): mini-batch k-means with EMA
// centroids in `passes` shuffled passes, with mass-based revival.
Clusters emeans(const float* x, uint64_t n, uint64_t d, uint64_t k, uint64_t passes, std::optional<uint32_t> seed,
                bool spherical, const NearestFn& nearest);

// PQCodebook::train (src/pq.cpp:41-73): k-means of each subspace (iters
// iterations, seed + s) over a seeded sample of at most 256 * K residuals.
// Returns the m x K x (dim / m) codebook.
std::vector<float> pq_codebooks(const float* residuals, uint64_t n, uint64_t dim, uint64_t m, uint64_t K,
                                uint64_t iters, std::optional<uint32_t> seed, const SubspaceNearestFn& nearest);

// Squared L2 distance in l2_distance_avx2's summation order (src/distance.cpp:69-101).
float l2_avx2(const float* a, const float* b, uint64_t n);

// detail::balanced_pca_rotation (src/rotation.cpp:71-130): the sample's
// covariance in double, its eigenvectors by cyclic Jacobi sweeps
// (symmetric_eigen, :20-69), and each eigenvector, largest eigenvalue first,
// placed in the subspace of room whose log-eigenvalue sum is smallest. Returns
// the d x d row-major R (x R rotates x). With the reference's arithmetic, as
// its g++ -O3 -mavx2 -mfma build fuses it, so R holds the same floats.
std::vector<float> balanced_pca_rotation(const float* x, uint64_t n, uint64_t d, uint64_t m);

} // namespace qh
