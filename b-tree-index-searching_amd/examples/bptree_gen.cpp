// Seeded generators for the reference's raw uint64 input files.
//
//   bptree_gen keys    <n> <file> [seed] [max]
//       n distinct keys drawn uniformly from [1, max] (max default 8.9e11),
//       written in random order. Replaces randomgenator/main.cpp:29-56, whose
//       reservoir loop over [n, 8.9e11) samples the same distribution.
//   bptree_gen queries <keyfile> <nq> <dist> [file] [seed]
//       nq point queries from the largest key M of <keyfile>; replaces
//       randomgenator/randgen.cu:58-246. dist:
//         0 uniform on [0, M]
//         1 normal, mean M/2, sd 0.35 * mean
//         2 Zipf, exponent 1, on [1, M]
//         3 five Gaussians, centres uniform on [0, M), sd 0.1, nq/5 each
//         4 normal, mean M/2, sd 0.63 * mean
//       Normal draws outside [0, M] are replaced by the mean (randgen.cu:123-126).
//       Without [file] the output name is the reference's
//       (exp-uniform-<keys/1e6>-Q<nq/1e6>, exp-normal-, zipf-, mulGaussian-,
//       normal-).
//   bptree_gen ranges  <keyfile> <nq> <dist> <startfile> <endfile> [seed] [log2 max width]
//       no reference counterpart: range queries whose starts follow <dist>
//       as above and whose widths are log-uniform on [1, 2^w] (w default 20),
//       end = start + width (saturating).
//
// The reference seeds from the clock or std::random_device, so its files are
// not reproducible; these are, for a given seed (default 0). Random numbers:
// xoshiro256** seeded through splitmix64 (Blackman & Vigna), normals by the
// Marsaglia polar method, Zipf by rejection-inversion (Hoermann & Derflinger,
// ACM TOMACS 6(3), 1996), which needs no table over [1, M] (the reference's
// randgen.cu:263-285 allocates M doubles).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

typedef unsigned long long ulli;

namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (uint64_t &w : s) {  // splitmix64
            x += 0x9e3779b97f4a7c15ull;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            w = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next() {  // xoshiro256**
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
    // uniform on [0, bound] without modulo bias
    uint64_t below_eq(uint64_t bound) {
        if (bound == ~0ull) return next();
        const uint64_t range = bound + 1, lim = ~0ull - (~0ull % range);
        uint64_t v;
        do v = next();
        while (v >= lim);
        return v % range;
    }
    double normal(double mean, double sd) {  // Marsaglia polar method
        if (have_spare) {
            have_spare = false;
            return mean + sd * spare;
        }
        double u, v, q;
        do {
            u = 2.0 * uniform() - 1.0;
            v = 2.0 * uniform() - 1.0;
            q = u * u + v * v;
        } while (q >= 1.0 || q == 0.0);
        const double f = std::sqrt(-2.0 * std::log(q) / q);
        spare = v * f;
        have_spare = true;
        return mean + sd * u * f;
    }
    bool have_spare = false;
    double spare = 0.0;
};

// Zipf on [1, n] with P(k) proportional to k^-s, s > 0, by rejection-inversion:
// H is an antiderivative of h(x) = x^-s; a continuous variate is drawn by
// inverting H between H(1.5) - h(1) and H(n + 0.5), rounded to the nearest
// integer k, and accepted when it lies close enough to k or under the hat.
struct Zipf {
    double s, n, hx1, hn, sq;
    Zipf(double exponent, double count) : s(exponent), n(count) {
        hx1 = H(1.5) - 1.0;
        hn = H(n + 0.5);
        sq = 2.0 - Hinv(H(2.5) - h(2.0));
    }
    double h(double x) const { return std::exp(-s * std::log(x)); }
    double H(double x) const {
        const double lx = std::log(x);
        return std::fabs(1.0 - s) < 1e-12 ? lx : std::expm1((1.0 - s) * lx) / (1.0 - s);
    }
    double Hinv(double y) const {
        if (std::fabs(1.0 - s) < 1e-12) return std::exp(y);
        return std::exp(std::log1p(y * (1.0 - s)) / (1.0 - s));
    }
    uint64_t draw(Rng &r) const {
        for (;;) {
            const double u = hn + r.uniform() * (hx1 - hn);
            const double x = Hinv(u);
            double k = std::floor(x + 0.5);
            if (k < 1.0) k = 1.0;
            if (k > n) k = n;
            if (k - x <= sq || u >= H(k + 0.5) - h(k)) return (uint64_t)k;
        }
    }
};

bool write_file(const std::string &path, const std::vector<ulli> &v) {
    std::ofstream f(path, std::ios::binary);
    if (!f) {
        fprintf(stderr, "Error: cannot write %s\n", path.c_str());
        return false;
    }
    f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(ulli)));
    return (bool)f;
}

bool read_file(const std::string &path, std::vector<ulli> &v) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) {
        fprintf(stderr, "Error: cannot open %s\n", path.c_str());
        return false;
    }
    const std::streamsize bytes = f.tellg();
    f.seekg(0);
    v.resize((size_t)bytes / sizeof(ulli));
    f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(v.size() * sizeof(ulli)));
    return (bool)f;
}

// n distinct keys uniform on [1, max], in random order.
bool gen_keys(ulli n, ulli max, uint64_t seed, std::vector<ulli> &out) {
    if (max < n) {
        fprintf(stderr, "Error: cannot draw %llu distinct keys from [1, %llu]\n", n, max);
        return false;
    }
    Rng r(seed);
    std::vector<ulli> k;
    k.reserve(n + n / 16 + 16);
    while (k.size() < n) {
        const ulli want = n - k.size();
        for (ulli i = 0; i < want + want / 16 + 16; ++i) k.push_back(1 + r.below_eq(max - 1));
        std::sort(k.begin(), k.end());
        k.erase(std::unique(k.begin(), k.end()), k.end());
    }
    // a uniform n-subset (partial Fisher-Yates), then a uniform order
    for (ulli i = 0; i < n; ++i) std::swap(k[i], k[i + r.below_eq(k.size() - 1 - i)]);
    k.resize(n);
    for (ulli i = n; i > 1; --i) std::swap(k[i - 1], k[r.below_eq(i - 1)]);
    out.swap(k);
    return true;
}

// One draw of distribution `dist` (0..4) on [0, M]; `centre` is the Gaussian
// centre of the current fifth for dist 3.
ulli draw(int dist, ulli M, Rng &r, const Zipf *z, ulli centre) {
    const double mean = (double)(M / 2);
    auto clamp = [&](double x, ulli fallback) -> ulli {
        return (x < 0.0 || x > (double)M || x >= 0x1.0p64) ? fallback : (ulli)x;
    };
    switch (dist) {
        case 0: return r.below_eq(M);
        case 1: return clamp(r.normal(mean, 0.35 * mean), M / 2);
        case 2: return z->draw(r);
        case 3: return clamp(r.normal((double)centre, 0.1), centre);
        default: return clamp(r.normal(mean, 0.63 * mean), M / 2);
    }
}

bool gen_queries(const std::vector<ulli> &keys, ulli nq, int dist, uint64_t seed, std::vector<ulli> &out) {
    if (keys.empty() || dist < 0 || dist > 4) {
        fprintf(stderr, "Error: need a non-empty key file and dist in 0..4\n");
        return false;
    }
    const ulli M = *std::max_element(keys.begin(), keys.end());
    Rng r(seed);
    Zipf z(1.0, (double)(M ? M : 1));
    out.resize(nq);
    const ulli per = nq / 5 ? nq / 5 : 1;
    ulli centre = 0;
    for (ulli i = 0; i < nq; ++i) {
        if (dist == 3 && i % per == 0 && i / per < 5) centre = M ? r.below_eq(M - 1) : 0;
        out[i] = draw(dist, M, r, &z, centre);
    }
    return true;
}

const char *ref_name(int dist) {
    static const char *names[5] = {"exp-uniform-", "exp-normal-", "zipf-", "mulGaussian-", "normal-"};
    return names[dist];
}

int usage(const char *prog) {
    fprintf(stderr,
            "usage: %s keys <n> <file> [seed] [max]\n"
            "       %s queries <keyfile> <nq> <dist 0..4> [file] [seed]\n"
            "       %s ranges <keyfile> <nq> <dist 0..4> <startfile> <endfile> [seed] [log2 max width]\n",
            prog, prog, prog);
    return -1;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) return usage(argv[0]);
    const std::string cmd = argv[1];
    if (cmd == "keys" && argc >= 4) {
        const ulli n = strtoull(argv[2], nullptr, 10);
        const uint64_t seed = argc > 4 ? strtoull(argv[4], nullptr, 10) : 0;
        const ulli max = argc > 5 ? strtoull(argv[5], nullptr, 10) : 890000000000ull;
        std::vector<ulli> k;
        if (!gen_keys(n, max, seed, k) || !write_file(argv[3], k)) return -1;
        printf("File saved: %s (%llu keys in [1, %llu])\n", argv[3], n, max);
        return 0;
    }
    if (cmd == "queries" && argc >= 5) {
        std::vector<ulli> keys, q;
        if (!read_file(argv[2], keys)) return -1;
        const ulli nq = strtoull(argv[3], nullptr, 10);
        const int dist = atoi(argv[4]);
        const uint64_t seed = argc > 6 ? strtoull(argv[6], nullptr, 10) : 0;
        if (!gen_queries(keys, nq, dist, seed, q)) return -1;
        const std::string file = argc > 5 ? std::string(argv[5])
                                          : std::string(ref_name(dist)) + std::to_string(keys.size() / 1000000) +
                                                "-Q" + std::to_string(nq / 1000000);
        if (!write_file(file, q)) return -1;
        printf("File saved: %s\n", file.c_str());
        return 0;
    }
    if (cmd == "ranges" && argc >= 7) {
        std::vector<ulli> keys, s;
        if (!read_file(argv[2], keys)) return -1;
        const ulli nq = strtoull(argv[3], nullptr, 10);
        const int dist = atoi(argv[4]);
        const uint64_t seed = argc > 7 ? strtoull(argv[7], nullptr, 10) : 0;
        const int w = argc > 8 ? atoi(argv[8]) : 20;
        if (w < 0 || w > 63 || !gen_queries(keys, nq, dist, seed, s)) return -1;
        Rng r(seed ^ 0x5bd1e995ull);
        std::vector<ulli> e(nq);
        for (ulli i = 0; i < nq; ++i) {
            const ulli width = (ulli)std::floor(std::exp2(r.uniform() * w));
            e[i] = s[i] > ~0ull - width ? ~0ull : s[i] + width;
        }
        if (!write_file(argv[5], s) || !write_file(argv[6], e)) return -1;
        printf("Files saved: %s %s\n", argv[5], argv[6]);
        return 0;
    }
    return usage(argv[0]);
}
