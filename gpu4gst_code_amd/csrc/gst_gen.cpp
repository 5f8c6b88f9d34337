// Deterministic synthetic instances: random connected graphs, Barabasi-Albert, R-MAT, groups.
// All randomness comes from the library's own generators (splitmix64 for seeding and
// counter-based streams, xoshiro256** for sequential streams), with rejection sampling for
// bounded integers and Box-Muller for normals, so one seed names one instance everywhere.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <unordered_set>
#include <vector>

#include "../../include/gst.h"

namespace {

inline uint64_t splitmix64_next(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

struct Xoshiro256 {
    uint64_t s[4];
    explicit Xoshiro256(uint64_t seed) {
        uint64_t sm = seed;
        for (int i = 0; i < 4; ++i) s[i] = splitmix64_next(sm);
    }
    uint64_t next() {
        const uint64_t result = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return result;
    }
    // uniform integer in [0, range) by rejection of the biased low band
    uint64_t below(uint64_t range) {
        if (range <= 1) return 0;
        const uint64_t threshold = (0 - range) % range;
        uint64_t x;
        do { x = next(); } while (x < threshold);
        return x % range;
    }
    int64_t uniform_int(int64_t lo, int64_t hi) { return lo + (int64_t)below((uint64_t)(hi - lo + 1)); }
    double uniform01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {  // Box-Muller
        double u1;
        do { u1 = uniform01(); } while (u1 <= 0.0);
        const double u2 = uniform01();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct EdgeList {
    std::vector<int32_t> a, b;
    std::vector<uint32_t> w;  // int or float bits
};

// Build a symmetric CSR (rows sorted by neighbour id) from unique undirected edges.
gst_status build_csr(int32_t n, const EdgeList& el, int32_t w_type, gst_host_csr* out) {
    const int64_t ne = (int64_t)el.a.size();
    std::vector<int64_t> deg(n + 1, 0);
    for (int64_t i = 0; i < ne; ++i) { deg[el.a[i]]++; deg[el.b[i]]++; }
    int64_t* row = (int64_t*)std::malloc(sizeof(int64_t) * (size_t)(n + 1));
    int32_t* col = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(1, 2 * ne));
    uint32_t* w = (uint32_t*)std::malloc(sizeof(uint32_t) * (size_t)std::max<int64_t>(1, 2 * ne));
    if (!row || !col || !w) { std::free(row); std::free(col); std::free(w); return GST_OOM; }
    row[0] = 0;
    for (int32_t v = 0; v < n; ++v) row[v + 1] = row[v] + deg[v];
    std::vector<int64_t> fill(row, row + n);
    for (int64_t i = 0; i < ne; ++i) {
        const int32_t x = el.a[i], y = el.b[i];
        col[fill[x]] = y; w[fill[x]] = el.w[i]; fill[x]++;
        col[fill[y]] = x; w[fill[y]] = el.w[i]; fill[y]++;
    }
    // sort each row by neighbour id (rows are short; insertion sort for small, std::sort else)
    std::vector<std::pair<int32_t, uint32_t>> tmp;
    for (int32_t v = 0; v < n; ++v) {
        const int64_t s = row[v], e = row[v + 1];
        if (e - s < 2) continue;
        tmp.resize((size_t)(e - s));
        for (int64_t k = s; k < e; ++k) tmp[(size_t)(k - s)] = {col[k], w[k]};
        std::sort(tmp.begin(), tmp.end(),
                  [](const std::pair<int32_t, uint32_t>& p, const std::pair<int32_t, uint32_t>& q) {
                      return p.first < q.first;
                  });
        for (int64_t k = s; k < e; ++k) { col[k] = tmp[(size_t)(k - s)].first; w[k] = tmp[(size_t)(k - s)].second; }
    }
    out->n = n; out->m = 2 * ne; out->row_ptr = row; out->col = col; out->w = w; out->w_type = w_type;
    return GST_OK;
}

// LSD radix sort of 64-bit keys (used for R-MAT edge deduplication)
void radix_sort_u64(std::vector<uint64_t>& keys, int key_bits) {
    std::vector<uint64_t> tmp(keys.size());
    const int digit = 11;
    const size_t buckets = (size_t)1 << digit;
    std::vector<size_t> cnt(buckets);
    for (int shift = 0; shift < key_bits; shift += digit) {
        std::fill(cnt.begin(), cnt.end(), 0);
        for (uint64_t k : keys) cnt[(k >> shift) & (buckets - 1)]++;
        size_t sum = 0;
        for (size_t i = 0; i < buckets; ++i) { size_t c = cnt[i]; cnt[i] = sum; sum += c; }
        for (uint64_t k : keys) tmp[cnt[(k >> shift) & (buckets - 1)]++] = k;
        keys.swap(tmp);
    }
}

}  // namespace

extern "C" {

void gst_host_csr_free(gst_host_csr* h) {
    if (!h) return;
    std::free(h->row_ptr); std::free(h->col); std::free(h->w);
    h->row_ptr = nullptr; h->col = nullptr; h->w = nullptr; h->n = 0; h->m = 0;
}

gst_status gst_gen_random_connected(int32_t n, int64_t n_edges, int32_t wlo, int32_t whi,
                                    uint64_t seed, gst_host_csr* out) {
    if (!out || n < 1 || wlo < 0 || whi < wlo) return GST_INVALID_ARG;
    const int64_t max_e = (int64_t)n * (n - 1) / 2;
    if (n_edges < n - 1 || n_edges > max_e) return GST_INVALID_ARG;
    Xoshiro256 rng(seed);
    EdgeList el;
    el.a.reserve((size_t)n_edges); el.b.reserve((size_t)n_edges); el.w.reserve((size_t)n_edges);
    std::unordered_set<uint64_t> seen;
    seen.reserve((size_t)n_edges * 2);
    auto key = [n](int32_t x, int32_t y) {
        if (x > y) std::swap(x, y);
        return (uint64_t)x * (uint64_t)n + (uint64_t)y;
    };
    // random spanning tree: vertex i joins a uniform earlier vertex
    for (int32_t i = 1; i < n; ++i) {
        const int32_t j = (int32_t)rng.below((uint64_t)i);
        el.a.push_back(j); el.b.push_back(i);
        el.w.push_back((uint32_t)rng.uniform_int(wlo, whi));
        seen.insert(key(i, j));
    }
    // uniform extra edges, rejecting self-loops and duplicates
    while ((int64_t)el.a.size() < n_edges) {
        const int32_t x = (int32_t)rng.below((uint64_t)n);
        const int32_t y = (int32_t)rng.below((uint64_t)n);
        if (x == y) continue;
        if (!seen.insert(key(x, y)).second) continue;
        el.a.push_back(x); el.b.push_back(y);
        el.w.push_back((uint32_t)rng.uniform_int(wlo, whi));
    }
    return build_csr(n, el, GST_W_INT32, out);
}

gst_status gst_gen_barabasi_albert(int32_t n, int32_t m, int32_t wlo, int32_t whi,
                                   uint64_t seed, gst_host_csr* out) {
    if (!out || m < 1 || n <= m || wlo < 0 || whi < wlo) return GST_INVALID_ARG;
    Xoshiro256 rng(seed);
    EdgeList el;
    const int64_t ne = (int64_t)m * (m - 1) / 2 + (int64_t)(n - m) * m;
    el.a.reserve((size_t)ne); el.b.reserve((size_t)ne); el.w.reserve((size_t)ne);
    std::vector<int32_t> ends;  // every edge endpoint once: sampling it is degree-proportional
    ends.reserve((size_t)(2 * ne));
    for (int32_t i = 0; i < m; ++i)
        for (int32_t j = i + 1; j < m; ++j) {
            el.a.push_back(i); el.b.push_back(j);
            el.w.push_back((uint32_t)rng.uniform_int(wlo, whi));
            ends.push_back(i); ends.push_back(j);
        }
    std::vector<int32_t> picked;
    for (int32_t v = m; v < n; ++v) {
        picked.clear();
        while ((int32_t)picked.size() < m) {
            int32_t t;
            if (ends.empty()) t = (int32_t)rng.below((uint64_t)v);  // m == 1 seed has no edges
            else t = ends[(size_t)rng.below(ends.size())];
            if (std::find(picked.begin(), picked.end(), t) != picked.end()) continue;
            picked.push_back(t);
        }
        for (int32_t t : picked) {
            el.a.push_back(t); el.b.push_back(v);
            el.w.push_back((uint32_t)rng.uniform_int(wlo, whi));
            ends.push_back(t); ends.push_back(v);
        }
    }
    return build_csr(n, el, GST_W_INT32, out);
}

gst_status gst_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                        int32_t w_type, int32_t wlo, int32_t whi, double mu, double sigma,
                        uint64_t seed, gst_host_csr* out) {
    if (!out || scale < 1 || scale > 30 || edge_factor < 1) return GST_INVALID_ARG;
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return GST_INVALID_ARG;
    if (w_type == GST_W_INT32 && (wlo < 0 || whi < wlo)) return GST_INVALID_ARG;
    if (w_type != GST_W_INT32 && w_type != GST_W_FLOAT32) return GST_INVALID_ARG;
    const int32_t n = (int32_t)1 << scale;
    const int64_t ne = (int64_t)edge_factor << scale;
    // vertex relabelling (Fisher-Yates) so hubs are spread over the id space
    std::vector<int32_t> perm((size_t)n);
    for (int32_t i = 0; i < n; ++i) perm[(size_t)i] = i;
    {
        Xoshiro256 prng(seed ^ 0x5bd1e995u);
        for (int32_t i = n - 1; i > 0; --i) {
            const int32_t j = (int32_t)prng.below((uint64_t)i + 1);
            std::swap(perm[(size_t)i], perm[(size_t)j]);
        }
    }
    const double ab = a + b, abc = a + b + c;
    std::vector<uint64_t> keys;
    keys.reserve((size_t)ne);
    for (int64_t e = 0; e < ne; ++e) {
        // counter-based stream per edge
        uint64_t st = seed * 0xD1B54A32D192ED03ull + (uint64_t)e * 0x9E3779B97F4A7C15ull;
        st = splitmix64_next(st);
        int32_t r = 0, q = 0;
        for (int lvl = 0; lvl < scale; ++lvl) {
            const double u = (double)(splitmix64_next(st) >> 11) * (1.0 / 9007199254740992.0);
            int rb, cb;
            if (u < a) { rb = 0; cb = 0; }
            else if (u < ab) { rb = 0; cb = 1; }
            else if (u < abc) { rb = 1; cb = 0; }
            else { rb = 1; cb = 1; }
            r = (r << 1) | rb; q = (q << 1) | cb;
        }
        int32_t x = perm[(size_t)r], y = perm[(size_t)q];
        if (x == y) continue;
        if (x > y) std::swap(x, y);
        keys.push_back(((uint64_t)x << 32) | (uint64_t)y);
    }
    radix_sort_u64(keys, 32 + scale);
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    EdgeList el;
    el.a.resize(keys.size()); el.b.resize(keys.size()); el.w.resize(keys.size());
    Xoshiro256 wrng(seed ^ 0xA5A5A5A5DEADBEEFull);
    for (size_t i = 0; i < keys.size(); ++i) {
        el.a[i] = (int32_t)(keys[i] >> 32);
        el.b[i] = (int32_t)(keys[i] & 0xFFFFFFFFu);
        if (w_type == GST_W_INT32) {
            el.w[i] = (uint32_t)wrng.uniform_int(wlo, whi);
        } else {
            const float f = (float)std::exp(mu + sigma * wrng.normal());
            uint32_t bits; std::memcpy(&bits, &f, 4);
            el.w[i] = bits;
        }
    }
    keys.clear(); keys.shrink_to_fit();
    return build_csr(n, el, w_type, out);
}

gst_status gst_gen_groups(int32_t n, int32_t g, int32_t smin, int32_t smax, uint64_t seed,
                          int32_t* grp_ptr, int32_t* grp_vtx) {
    if (n < 1 || g < 1 || smin < 1 || smax < smin || !grp_ptr || !grp_vtx) return GST_INVALID_ARG;
    Xoshiro256 rng(seed);
    grp_ptr[0] = 0;
    std::unordered_set<int32_t> inside;
    for (int32_t i = 0; i < g; ++i) {
        int32_t sz = (int32_t)rng.uniform_int(smin, smax);
        if (sz > n) sz = n;
        inside.clear();
        int32_t k = grp_ptr[i];
        if ((int64_t)sz * 2 > n) {
            // dense draw: partial Fisher-Yates over all vertices
            std::vector<int32_t> all((size_t)n);
            for (int32_t v = 0; v < n; ++v) all[(size_t)v] = v;
            for (int32_t t = 0; t < sz; ++t) {
                const int32_t j = t + (int32_t)rng.below((uint64_t)(n - t));
                std::swap(all[(size_t)t], all[(size_t)j]);
                grp_vtx[k++] = all[(size_t)t];
            }
        } else {
            while ((int32_t)inside.size() < sz) {
                const int32_t v = (int32_t)rng.below((uint64_t)n);
                if (inside.insert(v).second) grp_vtx[k++] = v;
            }
        }
        grp_ptr[i + 1] = k;
    }
    return GST_OK;
}

}  // extern "C"
