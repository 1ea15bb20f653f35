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

// Incoming table: for every target j the pairs {source, weight} in ascending
// source order.  Thread k owns the source range R_k; counts per (k, j) give
// each thread its write cursor, so the fill is parallel and the order within
// a target is ascending because ranges are ascending in k.
void build_incoming(int64_t n, int32_t nsyn, const uint2 *tab, int64_t *in_ptr, uint2 *in_tab, int threads)
{
    const int T = std::min<int64_t>(host_threads(threads), std::max<int64_t>(1, n / 1024));
    const int64_t per = (n + T - 1) / T;
    std::vector<std::vector<int64_t>> cur((size_t)T, std::vector<int64_t>((size_t)n, 0));
    {
        std::vector<std::thread> pool;
        for (int k = 0; k < T; ++k)
            pool.emplace_back([&, k] {
                const int64_t lo = k * per, hi = std::min<int64_t>(n, lo + per);
                auto &c = cur[(size_t)k];
                for (int64_t i = lo; i < hi; ++i) {
                    const uint2 *row = tab + (size_t)i * (size_t)nsyn;
                    for (int32_t e = 0; e < nsyn; ++e) c[row[e].x] += 1;
                }
            });
        for (auto &th : pool) th.join();
    }
    int64_t acc = 0;
    for (int64_t j = 0; j < n; ++j) {
        in_ptr[j] = acc;
        for (int k = 0; k < T; ++k) {
            const int64_t c = cur[(size_t)k][(size_t)j];
            cur[(size_t)k][(size_t)j] = acc;
            acc += c;
        }
    }
    in_ptr[n] = acc;
    {
        std::vector<std::thread> pool;
        for (int k = 0; k < T; ++k)
            pool.emplace_back([&, k] {
                const int64_t lo = k * per, hi = std::min<int64_t>(n, lo + per);
                auto &c = cur[(size_t)k];
                for (int64_t i = lo; i < hi; ++i) {
                    const uint2 *row = tab + (size_t)i * (size_t)nsyn;
                    for (int32_t e = 0; e < nsyn; ++e) {
                        const uint32_t j = row[e].x;
                        in_tab[(size_t)c[j]++] = make_uint2((uint32_t)i, row[e].y);
                    }
                }
            });
        for (auto &th : pool) th.join();
    }
}

}  // namespace g4s
