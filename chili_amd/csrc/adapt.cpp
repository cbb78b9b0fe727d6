// adapt.cpp -- host-side grid and channel-weight adaptation (tiny, d x B and
// C numbers per iteration) and the Philox helper of the C-ABI.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "common.hpp"

extern "C" {

// AdaptiveMap::Adapt, src/Integrator/AdaptiveMap.cc:67-125, on one grid of
// dims x (bins+1) edges with histogram data[dims][bins].  A dimension whose
// smoothed histogram is all zero keeps the reference's result: every edge
// of that dimension becomes 0.  The boundary scan is bounded by bins (the
// reference loop at AdaptiveMap.cc:112 is not).
int chili_grid_adapt(int dims, int bins, double* edges, double alpha, const double* data) {
    return guarded([&] {
        if (dims <= 0 || bins < 2 || !edges || !data) throw chili_error(CHILI_EINVAL, "bad grid");
        const size_t B = (size_t)bins, W = B + 1;
        std::vector<double> nh((size_t)dims * W, 0.0), sm(B);
        for (int i = 0; i < dims; ++i) {
            const double* row = data + (size_t)i * B;
            // three-point smoothing, end bins averaged over two
            double prev = row[0], cur = row[1];
            sm[0] = (prev + cur) / 2;
            double norm = sm[0];
            for (size_t b = 1; b + 1 < B; ++b) {
                const double s = prev + cur;
                prev = cur;
                cur = row[b + 1];
                sm[b] = (s + cur) / 3;
                norm += sm[b];
            }
            sm[B - 1] = (prev + cur) / 2;
            norm += sm[B - 1];
            if (norm == 0) continue;
            double avg = 0;
            for (size_t b = 0; b < B; ++b) {
                if (sm[b] != 0) {
                    const double r = sm[b] / norm;
                    const double fac = std::pow((r - 1.0) / std::log(r), alpha);
                    avg += fac;
                    sm[b] = fac;
                }
            }
            avg /= (double)B;
            const double* e = edges + (size_t)i * W;
            double acc = 0;
            size_t ib = 0;
            for (size_t nb = 1; nb < B; ++nb) {
                while (acc < avg && ib < B) acc += sm[ib++];
                const double lo = e[ib - 1], hi = e[ib];
                acc -= avg;
                nh[(size_t)i * W + nb] = hi - (hi - lo) * acc / sm[ib - 1];
            }
            nh[(size_t)i * W + B] = 1.0;
        }
        std::copy(nh.begin(), nh.end(), edges);
    });
}

// AdaptiveMap::Split, src/Integrator/AdaptiveMap.cc:127-160: every bin into
// nsplit equal parts; out holds dims x (nsplit*bins + 1) edges.
int chili_grid_split(int dims, int bins, const double* edges, int nsplit, double* out) {
    return guarded([&] {
        if (nsplit < 2 || nsplit > 4) throw chili_error(CHILI_EINVAL, "Invalid histogram splitting");
        const size_t W = (size_t)bins + 1, NW = (size_t)nsplit * bins + 1;
        for (int d = 0; d < dims; ++d) {
            for (int b = 0; b < bins; ++b) {
                const double lo = edges[d * W + b];
                const double size = edges[d * W + b + 1] - lo;
                for (int i = 0; i < nsplit; ++i) out[d * NW + (size_t)b * nsplit + i] = lo + (double)i * size / (double)nsplit;
            }
            out[d * NW + NW - 1] = 1.0;
        }
    });
}

// MultiChannel::Adapt, src/Channel/MultiChannel.cc:109-136:
// alpha_i <- alpha_i T_i^beta, normalised, clamped below at min_alpha (zeros
// stay zero), renormalised.
int chili_alpha_adapt(int n, double* alpha, const double* train, double beta, double min_alpha) {
    return guarded([&] {
        if (n <= 0 || !alpha || !train) throw chili_error(CHILI_EINVAL, "bad arguments");
        std::vector<double> w(n);
        double sum = 0;
        for (int i = 0; i < n; ++i) { w[i] = alpha[i] * std::pow(train[i], beta); sum += w[i]; }
        double ns = 0;
        for (auto& x : w) {
            if (x == 0) continue;
            x /= sum;
            x = std::max(x, min_alpha);
            ns += x;
        }
        for (int i = 0; i < n; ++i) alpha[i] = w[i] / ns;
    });
}

// Largest |T_i - T_j| (MultiChannel::MaxDifference, src/Channel/MultiChannel.cc:138-154).
double chili_max_difference(int n, const double* train) {
    double mx = 0;
    for (int i = 0; i + 1 < n; ++i)
        for (int j = i + 1; j < n; ++j) mx = std::max(mx, std::abs(train[i] - train[j]));
    return mx;
}

// Philox4x32-10 uniform as the kernels draw it: counter (attempt lo, attempt
// hi, k / 2, stream), key = seed, 53-bit double from the word pair of k.
double chili_philox_uniform(uint64_t seed, uint32_t stream, uint64_t attempt, uint32_t k) {
    uint32_t c0 = (uint32_t)attempt, c1 = (uint32_t)(attempt >> 32), c2 = k >> 1, c3 = stream;
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    const uint32_t hi = (k & 1u) ? c2 : c0, lo = (k & 1u) ? c3 : c1;
    return (double)(((uint64_t)hi << 21) | (uint64_t)(lo >> 11)) * 0x1p-53;
}

}  // extern "C"
