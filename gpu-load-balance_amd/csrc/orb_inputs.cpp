// orb_inputs.cpp -- synthetic particle sets for the benchmark workloads
// (host code).  No counterpart in the reference beyond its one generator
// (src/services/init.cu:11-25, 47-53), which orb_generate extends with a
// seed.
//
//  * orb_generate: the reference's xorshf96 stream, i-major draws, SoA.  Seed
//    0 gives the reference's seeds (init.cu:11); any other seed derives the
//    three state words from splitmix64 (Steele, Lea and Flood, "Fast
//    splittable pseudorandom number generators", OOPSLA 2014).
//  * orb_generate_plummer: a share of the particles from Plummer spheres of
//    scale radius a (centres uniform in [-0.5, 0.5)^3, equal split), the rest
//    uniform; every coordinate lies in [-0.5, 0.5) as float32.  Particle p's
//    draws come from its own counter-based stream (splitmix64 over a key of
//    the seed and p), and which sphere it belongs to from a keyed bijection of
//    [0, n) (a balanced Feistel network with cycle walking), so the set is
//    the same for any thread count and spheres are spread over the array.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include <sched.h>

#include "orb.h"

namespace {

inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

constexpr uint64_t kGolden = 0x9e3779b97f4a7c15ull;

// splitmix64: the state advances by the golden-ratio increment and each
// output is the mixed state.
struct SplitMix {
    uint64_t s;
    uint64_t next() { return mix64(s += kGolden); }
    // uniform in [0, 1) with 53 bits
    double unit() { return (double)(next() >> 11) * 0x1p-53; }
};

inline float xorshf96_step(uint64_t &x, uint64_t &y, uint64_t &z)
{
    // init.cu:15-24
    uint64_t t;
    x ^= x << 16;
    x ^= x >> 5;
    x ^= x << 1;
    t = x;
    x = y;
    y = z;
    z = t ^ x ^ y;
    float f = (float)z;
    float q = f / (float)UINT64_MAX;
    return (float)((double)q - 0.5);
}

int host_threads()
{
    cpu_set_t set;
    int n = 0;
    if (sched_getaffinity(0, sizeof set, &set) == 0) n = CPU_COUNT(&set);
    if (n < 1) n = (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(n, 64));
}

template <class F>
void parallel_ranges(int64_t n, F f)
{
    const int nt = (int)std::min<int64_t>(host_threads(), std::max<int64_t>(1, n / 65536));
    if (nt <= 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) {
        const int64_t a = n * t / nt, b = n * (t + 1) / nt;
        th.emplace_back([=] { f(a, b); });
    }
    for (auto &t : th) t.join();
}

// A bijection of [0, 2^(2h)) keyed by `key`: four Feistel rounds on two
// h-bit halves, each round function a keyed mix of the other half.
struct Feistel {
    int h;
    uint64_t mask;
    uint64_t keys[4];
    Feistel(int64_t n, uint64_t key)
    {
        h = 1;
        while ((int64_t(1) << (2 * h)) < n) ++h;
        mask = (uint64_t(1) << h) - 1;
        SplitMix sm{key};
        for (auto &k : keys) k = sm.next();
    }
    uint64_t once(uint64_t v) const
    {
        uint64_t l = v >> h, r = v & mask;
        for (uint64_t k : keys) {
            const uint64_t t = l ^ (mix64(r ^ k) & mask);
            l = r;
            r = t;
        }
        return (l << h) | r;
    }
    // cycle walking: a bijection of [0, n) (the domain is below 4n)
    uint64_t operator()(uint64_t v, uint64_t n) const
    {
        do v = once(v);
        while (v >= n);
        return v;
    }
};

}  // namespace

extern "C" {

int orb_generate(uint64_t seed, int64_t n, float *x, float *y, float *z)
{
    if (n < 0 || (n > 0 && (!x || !y || !z))) return ORB_ERR_ARG;
    uint64_t sx = 123456789ull, sy = 362436069ull, sz = 521288629ull;  // init.cu:11
    if (seed != 0) {
        SplitMix sm{seed};
        do {
            sx = sm.next();
            sy = sm.next();
            sz = sm.next();
        } while ((sx | sy | sz) == 0);
    }
    for (int64_t i = 0; i < n; ++i) {  // init.cu:47-53: i-major draws
        x[i] = xorshf96_step(sx, sy, sz);
        y[i] = xorshf96_step(sx, sy, sz);
        z[i] = xorshf96_step(sx, sy, sz);
    }
    return ORB_OK;
}

int orb_generate_plummer(uint64_t seed, int64_t n, int32_t n_spheres, float a, float background,
                         float *x, float *y, float *z)
{
    if (n < 0 || (n > 0 && (!x || !y || !z)) || n_spheres < 1 || !(a > 0.0f) ||
        !(background >= 0.0f && background <= 1.0f) || n > (int64_t(1) << 40))
        return ORB_ERR_ARG;
    if (n == 0) return ORB_OK;
    const int64_t n_bg = (int64_t)std::llround((double)n * (double)background);
    const int64_t n_cl = n - n_bg;
    std::vector<double> centre(3 * (size_t)n_spheres);
    SplitMix cs{mix64(seed ^ 0x243f6a8885a308d3ull)};
    for (auto &c : centre) c = cs.unit() - 0.5;
    // equal split: the first `rem` spheres take one more
    const int64_t base = n_cl / n_spheres, rem = n_cl % n_spheres;
    const Feistel perm((int64_t)n, mix64(seed ^ 0x13198a2e03707344ull));
    const uint64_t pkey = mix64(seed ^ 0xa4093822299f31d0ull);
    const float top = std::nextafter(0.5f, 0.0f);
    const double two_pi = 6.283185307179586476925;
    parallel_ranges(n, [&](int64_t p0, int64_t p1) {
        for (int64_t p = p0; p < p1; ++p) {
            const int64_t j = (int64_t)perm((uint64_t)p, (uint64_t)n);
            SplitMix sm{mix64(pkey + (uint64_t)p * kGolden)};
            float v[3];
            if (j >= n_cl) {  // uniform background
                for (int d = 0; d < 3; ++d) v[d] = std::min((float)(sm.unit() - 0.5), top);
            } else {
                const int64_t s = j < rem * (base + 1) ? j / (base + 1)
                                                       : rem + (j - rem * (base + 1)) / base;
                const double *c = &centre[3 * (size_t)s];
                for (;;) {
                    // Plummer: enclosed mass fraction u = r^3 / (r^2 + a^2)^(3/2)
                    const double u = 1.0 - sm.unit();  // (0, 1]
                    const double r = (double)a / std::sqrt(std::pow(u, -2.0 / 3.0) - 1.0);
                    const double ct = 2.0 * sm.unit() - 1.0, ph = two_pi * sm.unit();
                    const double st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
                    v[0] = (float)(c[0] + r * st * std::cos(ph));
                    v[1] = (float)(c[1] + r * st * std::sin(ph));
                    v[2] = (float)(c[2] + r * ct);
                    bool in = std::isfinite(r);
                    for (int d = 0; d < 3; ++d) in = in && v[d] >= -0.5f && v[d] < 0.5f;
                    if (in) break;  // outside the box: drawn again
                }
            }
            x[p] = v[0];
            y[p] = v[1];
            z[p] = v[2];
        }
    });
    return ORB_OK;
}

}  // extern "C"
