// Synthetic graph generators and the CSR builder (host side, no GPU).
//
// Every generator draws arc i (or row r, or vertex v) from a random stream
// that depends only on (seed, i >> 16) — so output is identical for any
// thread count.  The reference's own generators (MedusaRT/GraphGenerator.cu:21-209,
// rand32 in MedusaRT/Utilities.cu:50-62) are not reproduced: the workloads
// here are specified by their distributions, not by the reference's PRNG.

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <numeric>

#include "common.h"

namespace medusa {

static constexpr uint64_t kChunk = 1ull << 16;

}  // namespace medusa

using namespace medusa;

extern "C" int medusa_gen_erdos_renyi(uint32_t n, uint64_t m, uint64_t seed, uint64_t capacity,
                                      uint32_t* src, uint32_t* dst, uint64_t* m_out) {
    return guard([&] {
        MDS_CHECK(n > 0 && src && dst && m_out, MEDUSA_ERR_ARG, "gen_erdos_renyi: bad arguments");
        MDS_CHECK(capacity >= m, MEDUSA_ERR_ARG, "gen_erdos_renyi: capacity %llu < m %llu",
                  (unsigned long long)capacity, (unsigned long long)m);
        const uint64_t chunks = (m + kChunk - 1) / kChunk;
        std::vector<uint64_t> kept(chunks + 1, 0);
        // draw every chunk into place, compacting self-loops out within the chunk
        parallel_blocks(chunks, [&](uint64_t cb, uint64_t ce) {
            for (uint64_t c = cb; c < ce; ++c) {
                Rng rng(seed, c);
                const uint64_t b = c * kChunk, e = std::min(m, b + kChunk);
                uint64_t w = b;
                for (uint64_t i = b; i < e; ++i) {
                    const uint32_t u = rng.below(n), v = rng.below(n);
                    if (u == v) continue;
                    src[w] = u;
                    dst[w] = v;
                    ++w;
                }
                kept[c + 1] = w - b;
            }
        });
        // close the gaps between chunks, in order
        uint64_t out = 0;
        for (uint64_t c = 0; c < chunks; ++c) {
            const uint64_t b = c * kChunk, k = kept[c + 1];
            if (out != b) {
                std::memmove(src + out, src + b, k * sizeof(uint32_t));
                std::memmove(dst + out, dst + b, k * sizeof(uint32_t));
            }
            out += k;
        }
        *m_out = out;
    });
}

extern "C" int medusa_gen_rmat(uint32_t scale, uint64_t m, double a, double b, double c, double noise,
                               uint64_t seed, uint64_t capacity, uint32_t* src, uint32_t* dst,
                               uint64_t* m_out) {
    return guard([&] {
        MDS_CHECK(scale >= 1 && scale <= 31 && src && dst && m_out, MEDUSA_ERR_ARG, "gen_rmat: bad arguments");
        MDS_CHECK(capacity >= m, MEDUSA_ERR_ARG, "gen_rmat: capacity too small");
        const double d = 1.0 - a - b - c;
        MDS_CHECK(a >= 0 && b >= 0 && c >= 0 && d >= 0, MEDUSA_ERR_ARG, "gen_rmat: probabilities must sum to <= 1");
        MDS_CHECK(noise >= 0 && noise < 1, MEDUSA_ERR_ARG, "gen_rmat: noise must be in [0,1)");
        const uint64_t chunks = (m + kChunk - 1) / kChunk;
        parallel_blocks(chunks, [&](uint64_t cb, uint64_t ce) {
            for (uint64_t ch = cb; ch < ce; ++ch) {
                Rng rng(seed, ch);
                const uint64_t beg = ch * kChunk, end = std::min(m, beg + kChunk);
                for (uint64_t i = beg; i < end; ++i) {
                    uint32_t u = 0, v = 0;
                    for (uint32_t lvl = 0; lvl < scale; ++lvl) {
                        // per-level multiplicative noise on each quadrant, then renormalise
                        const double pa = a * (1.0 + noise * (2.0 * rng.uniform() - 1.0));
                        const double pb = b * (1.0 + noise * (2.0 * rng.uniform() - 1.0));
                        const double pc = c * (1.0 + noise * (2.0 * rng.uniform() - 1.0));
                        const double pd = d * (1.0 + noise * (2.0 * rng.uniform() - 1.0));
                        const double r = rng.uniform() * (pa + pb + pc + pd);
                        uint32_t row, colb;
                        if (r < pa) {
                            row = 0; colb = 0;
                        } else if (r < pa + pb) {
                            row = 0; colb = 1;
                        } else if (r < pa + pb + pc) {
                            row = 1; colb = 0;
                        } else {
                            row = 1; colb = 1;
                        }
                        u = (u << 1) | row;
                        v = (v << 1) | colb;
                    }
                    src[i] = u;
                    dst[i] = v;
                }
            }
        });
        *m_out = m;
    });
}

extern "C" int medusa_gen_lattice(uint32_t rows, uint32_t cols, double p_keep, double p_diag, uint32_t wmin,
                                  uint32_t wmax, uint64_t seed, uint64_t capacity, uint32_t* src,
                                  uint32_t* dst, uint32_t* weight, uint64_t* m_out) {
    return guard([&] {
        MDS_CHECK(rows >= 1 && cols >= 1 && src && dst && m_out, MEDUSA_ERR_ARG, "gen_lattice: bad arguments");
        MDS_CHECK((uint64_t)rows * cols < (1ull << 32), MEDUSA_ERR_ARG, "gen_lattice: too many vertices");
        MDS_CHECK(wmin >= 1 && wmax >= wmin, MEDUSA_ERR_ARG, "gen_lattice: need 1 <= wmin <= wmax");
        const uint32_t wspan = wmax - wmin + 1;  // 0 when the span is 2^32 (not allowed: wmin >= 1)
        // each row writes at most 6 arcs per cell (right, down, diagonal; both directions)
        std::vector<std::vector<uint32_t>> rs(rows), rd(rows), rw(rows);
        parallel_blocks(rows, [&](uint64_t rb, uint64_t re) {
            for (uint64_t r = rb; r < re; ++r) {
                Rng rng(seed, r);
                auto& S = rs[r];
                auto& D = rd[r];
                auto& W = rw[r];
                auto emit = [&](uint32_t x, uint32_t y, uint32_t w) {
                    S.push_back(x); D.push_back(y); W.push_back(w);
                    S.push_back(y); D.push_back(x); W.push_back(w);
                };
                for (uint32_t c = 0; c < cols; ++c) {
                    const uint32_t id = (uint32_t)r * cols + c;
                    if (c + 1 < cols && rng.uniform() < p_keep) emit(id, id + 1, wmin + rng.below(wspan));
                    if (r + 1 < rows) {
                        if (rng.uniform() < p_keep) emit(id, id + cols, wmin + rng.below(wspan));
                        if (rng.uniform() < p_diag) {
                            const bool left = (rng.next() >> 63) != 0;
                            const int64_t tc = left ? (int64_t)c - 1 : (int64_t)c + 1;
                            if (tc >= 0 && tc < (int64_t)cols)
                                emit(id, (uint32_t)((r + 1) * cols + tc), wmin + rng.below(wspan));
                        }
                    }
                }
            }
        });
        uint64_t total = 0;
        for (auto& v : rs) total += v.size();
        MDS_CHECK(capacity >= total, MEDUSA_ERR_ARG, "gen_lattice: capacity %llu < %llu arcs",
                  (unsigned long long)capacity, (unsigned long long)total);
        uint64_t o = 0;
        for (uint32_t r = 0; r < rows; ++r) {
            const size_t k = rs[r].size();
            std::memcpy(src + o, rs[r].data(), k * 4);
            std::memcpy(dst + o, rd[r].data(), k * 4);
            if (weight) std::memcpy(weight + o, rw[r].data(), k * 4);
            o += k;
        }
        *m_out = total;
    });
}

extern "C" int medusa_gen_chung_lu(uint32_t n, uint64_t m, double exponent, uint32_t max_degree,
                                   double p_zero, uint64_t seed, uint64_t capacity, uint32_t* src,
                                   uint32_t* dst, uint64_t* m_out) {
    return guard([&] {
        MDS_CHECK(n > 0 && max_degree >= 1 && src && dst && m_out, MEDUSA_ERR_ARG, "gen_chung_lu: bad arguments");
        MDS_CHECK(capacity >= m, MEDUSA_ERR_ARG, "gen_chung_lu: capacity too small");
        MDS_CHECK(p_zero >= 0 && p_zero < 1 && exponent > 0, MEDUSA_ERR_ARG, "gen_chung_lu: bad distribution");
        // Zipf(exponent) on [1, max_degree] by inverse CDF
        std::vector<double> cdf(max_degree);
        double acc = 0;
        for (uint32_t k = 1; k <= max_degree; ++k) {
            acc += std::pow((double)k, -exponent);
            cdf[k - 1] = acc;
        }
        // vertex weights: streams 2^40 + chunk keep them apart from the arc streams
        std::vector<uint64_t> wcum(n + 1, 0);
        const uint64_t vchunks = (n + kChunk - 1) / kChunk;
        parallel_blocks(vchunks, [&](uint64_t cb, uint64_t ce) {
            for (uint64_t ch = cb; ch < ce; ++ch) {
                Rng rng(seed, (1ull << 40) + ch);
                const uint64_t b = ch * kChunk, e = std::min<uint64_t>(n, b + kChunk);
                for (uint64_t v = b; v < e; ++v) {
                    uint64_t w = 0;
                    if (rng.uniform() >= p_zero) {
                        const double u = rng.uniform() * acc;
                        w = (uint64_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
                        if (w > max_degree) w = max_degree;
                    }
                    wcum[v + 1] = w;
                }
            }
        });
        for (uint32_t v = 0; v < n; ++v) wcum[v + 1] += wcum[v];
        const uint64_t W = wcum[n];
        MDS_CHECK(W > 0 || m == 0, MEDUSA_ERR_ARG, "gen_chung_lu: every vertex drew weight 0");
        const uint64_t chunks = (m + kChunk - 1) / kChunk;
        parallel_blocks(chunks, [&](uint64_t cb, uint64_t ce) {
            for (uint64_t ch = cb; ch < ce; ++ch) {
                Rng rng(seed, ch);
                const uint64_t b = ch * kChunk, e = std::min(m, b + kChunk);
                for (uint64_t i = b; i < e; ++i) {
                    const uint64_t x = rng.below64(W);
                    // first v with wcum[v+1] > x
                    const uint32_t u = (uint32_t)(std::upper_bound(wcum.begin() + 1, wcum.end(), x) - (wcum.begin() + 1));
                    src[i] = u;
                    dst[i] = rng.below(n);
                }
            }
        });
        *m_out = m;
    });
}

extern "C" int medusa_gen_weights(uint64_t m, uint32_t wmin, uint32_t wmax, uint64_t seed, uint32_t* weight) {
    return guard([&] {
        MDS_CHECK(weight || m == 0, MEDUSA_ERR_ARG, "gen_weights: null output");
        MDS_CHECK(wmin >= 1 && wmax >= wmin, MEDUSA_ERR_ARG, "gen_weights: need 1 <= wmin <= wmax");
        const uint32_t span = wmax - wmin + 1;
        const uint64_t chunks = (m + kChunk - 1) / kChunk;
        parallel_blocks(chunks, [&](uint64_t cb, uint64_t ce) {
            for (uint64_t ch = cb; ch < ce; ++ch) {
                Rng rng(seed, (1ull << 41) + ch);
                const uint64_t b = ch * kChunk, e = std::min(m, b + kChunk);
                for (uint64_t i = b; i < e; ++i) weight[i] = wmin + rng.below(span);
            }
        });
    });
}

extern "C" int medusa_csr_from_arcs(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                                    const uint32_t* weight, uint32_t* row_ptr, uint32_t* col,
                                    uint32_t* weight_out) {
    return guard([&] {
        MDS_CHECK(row_ptr && (m == 0 || (src && dst && col)), MEDUSA_ERR_ARG, "csr_from_arcs: null argument");
        MDS_CHECK(m < (1ull << 32), MEDUSA_ERR_ARG, "csr_from_arcs: m must be < 2^32");
        MDS_CHECK(!weight_out || weight, MEDUSA_ERR_ARG, "csr_from_arcs: weight_out needs weight");
        std::vector<uint32_t> cnt(n + 1, 0);
        for (uint64_t i = 0; i < m; ++i) {
            MDS_CHECK(src[i] < n && dst[i] < n, MEDUSA_ERR_ARG, "csr_from_arcs: arc %llu (%u->%u) out of range n=%u",
                      (unsigned long long)i, src[i], dst[i], n);
            ++cnt[src[i] + 1];
        }
        for (uint32_t v = 0; v < n; ++v) cnt[v + 1] += cnt[v];
        std::memcpy(row_ptr, cnt.data(), (size_t)(n + 1) * 4);
        for (uint64_t i = 0; i < m; ++i) {
            const uint32_t pos = cnt[src[i]]++;
            col[pos] = dst[i];
            if (weight_out) weight_out[pos] = weight[i];
        }
    });
}
