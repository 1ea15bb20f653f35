// g4s_build.hip -- host-side connectivity builder (replaces the reference's
// O(N^2*Nsyn) sampler, cu:862-912) and the incoming-table transpose used by
// the ordered-gather mode.  Multithreaded C++, no device code.
#include "g4s_build.h"

#include <algorithm>
#include <thread>
#include <vector>

namespace g4s {

int host_threads(int requested)
{
    int t = requested > 0 ? requested : (int)std::thread::hardware_concurrency();
    if (t < 1) t = 1;
    if (t > 64) t = 64;
    return t;
}

template <class F>
static void parallel_rows(int64_t n, int threads, F &&f)
{
    if (threads <= 1 || n < 4096) {
        f(0, n, 0);
        return;
    }
    std::vector<std::thread> pool;
    const int64_t per = (n + threads - 1) / threads;
    for (int k = 0; k < threads; ++k) {
        const int64_t lo = k * per, hi = std::min<int64_t>(n, lo + per);
        if (lo >= hi) break;
        pool.emplace_back([&, lo, hi, k] { f(lo, hi, k); });
    }
    for (auto &th : pool) th.join();
}

// Row i: Floyd's algorithm draws Nsyn distinct values from the m = N-1
// non-self candidates (for jj = m-Nsyn .. m-1: t uniform on [0, jj] by a
// 64x64->128 multiply-shift of Philox words 0,1 under counter (i, d, CONN);
// insert t, or jj if t is present); value v maps to neuron v + (v >= i); the
// row is sorted; weight of (i -> j) = g * U(Philox(i, j, WEIGHT).x0) with
// g = 0.5*1000/Nsyn (excitatory rows) or -1000/Nsyn (inhibitory rows),
// cu:890, cu:893.
void build_table(uint64_t seed, int64_t n, int64_t n_exc, int32_t nsyn, uint2 *tab, int threads)
{
    const uint64_t m = (uint64_t)(n - 1);
    const float scale = 1000.0f / (float)nsyn;
    const float g_exc = scale * 0.5f, g_inh = -scale;
    parallel_rows(n, host_threads(threads), [&](int64_t lo, int64_t hi, int) {
        std::vector<uint8_t> mark((size_t)m, 0);
        std::vector<uint32_t> row((size_t)nsyn);
        for (int64_t i = lo; i < hi; ++i) {
            for (int32_t d = 0; d < nsyn; ++d) {
                const uint64_t jj = m - (uint64_t)nsyn + (uint64_t)d;
                const Word4 o = draw(seed, (uint32_t)i, (uint32_t)d, kTagConn);
                const uint64_t r = ((uint64_t)o.x1 << 32) | o.x0;
                const uint64_t t = (uint64_t)(((unsigned __int128)r * (jj + 1u)) >> 64);
                const uint64_t val = mark[t] ? jj : t;
                mark[val] = 1;
                row[d] = (uint32_t)val;
            }
            for (int32_t d = 0; d < nsyn; ++d) {
                mark[row[d]] = 0;
                if ((int64_t)row[d] >= i) row[d] += 1u;
            }
            std::sort(row.begin(), row.end());
            const float g = (i < n_exc) ? g_exc : g_inh;
            uint2 *out = tab + (size_t)i * (size_t)nsyn;
            for (int32_t d = 0; d < nsyn; ++d) {
                const Word4 o = draw(seed, (uint32_t)i, row[d], kTagWeight);
                out[d] = make_uint2(row[d], f2u(g * u01(o.x0)));
            }
        }
    });
}

}  // namespace g4s
