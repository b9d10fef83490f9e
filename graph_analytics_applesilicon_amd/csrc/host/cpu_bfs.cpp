// CPU BFS variants and the hybrid BFS's CPU level steps.
//   bfs_serial  src/cpu/bfs_cpu.cpp:27-45   (FIFO BFS, the bit-exact oracle of gbench --validate)
//   bfs_omp     src/cpu/bfs_cpu.cpp:207-223 (level-synchronous top-down, runtime schedule)
//   bfs_do      src/cpu/bfs_cpu.cpp:227-283 (direction-optimizing, GAP alpha/beta rule)
//   steps::*    src/cpu/bfs_steps.cpp:27-143, with u64 bitmap words (the GPU layout here)
#include "graph.hpp"

#include <omp.h>

#include <algorithm>
#include <cstring>

namespace ggpu {

namespace {

int nthreads(int t) { return t > 0 ? t : omp_get_max_threads(); }

double now_ms() { return omp_get_wtime() * 1e3; }

// First writer wins: a relaxed load filters visited vertices before the CAS.
inline bool claim(uint32_t* dist, uint32_t v, uint32_t d) {
    uint32_t seen = __atomic_load_n(&dist[v], __ATOMIC_RELAXED);
    if (seen != kInf) return false;
    return __atomic_compare_exchange_n(&dist[v], &seen, d, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED);
}

struct Level {
    uint64_t next = 0, scout = 0, examined = 0;
};

// Discoveries are buffered per thread and published with one atomic per block.
constexpr size_t kFlush = 4096;

Level td_step(const Csr& g, uint32_t* dist, const uint32_t* front, int64_t fsize, uint32_t* next, uint32_t d, int nt) {
    uint64_t tail = 0, scout = 0, examined = 0;
#pragma omp parallel num_threads(nt) reduction(+ : scout, examined)
    {
        std::vector<uint32_t> buf;
        buf.reserve(kFlush);
        auto publish = [&] {
            if (buf.empty()) return;
            const uint64_t at = __atomic_fetch_add(&tail, (uint64_t)buf.size(), __ATOMIC_RELAXED);
            std::memcpy(next + at, buf.data(), buf.size() * 4);
            buf.clear();
        };
#pragma omp for schedule(runtime) nowait
        for (int64_t i = 0; i < fsize; ++i) {
            const uint32_t u = front[i];
            examined += g.deg(u);
            for (const uint32_t* p = g.nbr_begin(u); p != g.nbr_end(u); ++p)
                if (claim(dist, *p, d)) {
                    scout += g.deg(*p);
                    buf.push_back(*p);
                    if (buf.size() == kFlush) publish();
                }
        }
        publish();
    }
    return {tail, scout, examined};
}

struct Bits {
    std::vector<uint64_t> w;
    explicit Bits(uint64_t n) : w((n + 63) / 64, 0) {}
    bool test(uint32_t i) const { return (w[i >> 6] >> (i & 63)) & 1; }
    void set_atomic(uint32_t i) { __atomic_fetch_or(&w[i >> 6], 1ull << (i & 63), __ATOMIC_RELAXED); }
    void zero(int nt) {
        const int64_t nw = (int64_t)w.size();
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int64_t i = 0; i < nw; ++i) w[i] = 0;
    }
};

// Bottom-up: each unvisited vertex stops at its first neighbour in the frontier.
Level bu_step(const Csr& g, uint32_t* dist, const Bits& front, Bits& next, uint32_t d, int nt) {
    uint64_t awake = 0, examined = 0;
    const int64_t n = (int64_t)g.n;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1024) reduction(+ : awake, examined)
    for (int64_t v = 0; v < n; ++v) {
        if (dist[v] != kInf) continue;
        const uint32_t* b = g.nbr_begin(v);
        const uint32_t* e = g.nbr_end(v);
        const uint32_t* p = b;
        for (; p != e; ++p)
            if (front.test(*p)) {
                dist[v] = d;
                next.set_atomic((uint32_t)v);
                ++awake;
                ++p;
                break;
            }
        examined += (uint64_t)(p - b);
    }
    return {awake, 0, examined};
}

void q2b(const uint32_t* q, int64_t size, Bits& bm, int nt) {
    bm.zero(nt);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t i = 0; i < size; ++i) bm.set_atomic(q[i]);
}

int64_t b2q(const Bits& bm, uint64_t n, uint32_t* q, int nt) {
    uint64_t tail = 0;
    const int64_t nw = (int64_t)bm.w.size();
#pragma omp parallel num_threads(nt)
    {
        std::vector<uint32_t> buf;
        buf.reserve(kFlush);
        auto publish = [&] {
            if (buf.empty()) return;
            const uint64_t at = __atomic_fetch_add(&tail, (uint64_t)buf.size(), __ATOMIC_RELAXED);
            std::memcpy(q + at, buf.data(), buf.size() * 4);
            buf.clear();
        };
#pragma omp for schedule(static) nowait
        for (int64_t i = 0; i < nw; ++i)
            for (uint64_t word = bm.w[i]; word; word &= word - 1) {
                const uint64_t v = (uint64_t)i * 64 + (uint64_t)__builtin_ctzll(word);
                if (v < n) {
                    buf.push_back((uint32_t)v);
                    if (buf.size() == kFlush) publish();
                }
            }
        publish();
    }
    return (int64_t)tail;
}

void set_schedule(Schedule s) {
    if (s == Schedule::Static) omp_set_schedule(omp_sched_static, 0);
    else if (s == Schedule::Dynamic64) omp_set_schedule(omp_sched_dynamic, 64);
    else omp_set_schedule(omp_sched_guided, 0);
}

ggpu_level make_level(uint32_t level, uint64_t frontier, uint64_t edges, char dir, double ms) {
    ggpu_level L{};
    L.level = level;
    L.action = dir == 'B' ? 2u : 0u;
    L.frontier = frontier;
    L.edges = edges;
    L.frontier_edges = dir == 'T' ? edges : 0;
    L.direction = dir;
    L.device = 'C';
    L.time_ms = ms;
    return L;
}

}  // namespace

void bfs_serial(const Csr& g, uint32_t src, uint32_t* dist) {
    std::fill(dist, dist + g.n, kInf);
    std::vector<uint32_t> fifo(g.n);
    uint64_t head = 0, tail = 0;
    dist[src] = 0;
    fifo[tail++] = src;
    while (head < tail) {
        const uint32_t u = fifo[head++];
        const uint32_t du = dist[u] + 1;
        for (const uint32_t* p = g.nbr_begin(u); p != g.nbr_end(u); ++p)
            if (dist[*p] == kInf) {
                dist[*p] = du;
                fifo[tail++] = *p;
            }
    }
}

void bfs_omp(const Csr& g, uint32_t src, uint32_t* dist, const CpuBfsOpts& o, std::vector<ggpu_level>* trace) {
    const int nt = nthreads(o.threads);
    set_schedule(o.schedule);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t v = 0; v < (int64_t)g.n; ++v) dist[v] = kInf;
    std::vector<uint32_t> front(g.n), next(g.n);
    dist[src] = 0;
    front[0] = src;
    int64_t fsize = 1;
    for (uint32_t level = 0; fsize > 0; ++level) {
        const double t0 = trace ? now_ms() : 0;
        const Level r = td_step(g, dist, front.data(), fsize, next.data(), level + 1, nt);
        if (trace) trace->push_back(make_level(level, (uint64_t)fsize, r.examined, 'T', now_ms() - t0));
        front.swap(next);
        fsize = (int64_t)r.next;
    }
}

void bfs_do(const Csr& g, uint32_t src, uint32_t* dist, const CpuBfsOpts& o, std::vector<ggpu_level>* trace) {
    const int nt = nthreads(o.threads);
    const uint64_t n = g.n;
    omp_set_schedule(omp_sched_dynamic, 64);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t v = 0; v < (int64_t)n; ++v) dist[v] = kInf;
    std::vector<uint32_t> queue(n), next_queue(n);
    Bits front(n), next(n);
    dist[src] = 0;
    queue[0] = src;
    int64_t fsize = 1;
    uint32_t level = 0;
    uint64_t to_check = g.m;      // m_u: edges not yet examined top-down
    uint64_t scout = g.deg(src);  // m_f: edges of the frontier
    while (fsize > 0) {
        if ((double)scout > (double)to_check / o.alpha) {
            double t0 = now_ms();
            q2b(queue.data(), fsize, front, nt);
            double conv = now_ms() - t0;
            uint64_t awake = (uint64_t)fsize, prev;
            do {  // bottom-up while the frontier grows or stays above n / beta
                prev = awake;
                const double t1 = now_ms();
                next.zero(nt);
                const Level r = bu_step(g, dist, front, next, level + 1, nt);
                if (trace) trace->push_back(make_level(level, prev, r.examined, 'B', now_ms() - t1 + conv));
                conv = 0;
                awake = r.next;
                std::swap(front.w, next.w);
                ++level;
            } while (awake >= prev || (double)awake > (double)n / o.beta);
            t0 = now_ms();
            fsize = b2q(front, n, queue.data(), nt);
            if (trace && !trace->empty()) trace->back().time_ms += now_ms() - t0;
            scout = 1;
        } else {
            const double t0 = trace ? now_ms() : 0;
            to_check -= std::min(to_check, scout);
            const Level r = td_step(g, dist, queue.data(), fsize, next_queue.data(), level + 1, nt);
            if (trace) trace->push_back(make_level(level, (uint64_t)fsize, r.examined, 'T', now_ms() - t0));
            scout = r.scout;
            queue.swap(next_queue);
            fsize = (int64_t)r.next;
            ++level;
        }
    }
}

namespace steps {

void init_dist(uint32_t* dist, uint64_t n, uint32_t src, int threads) {
#pragma omp parallel for num_threads(nthreads(threads)) schedule(static) if (n > (1u << 16))
    for (int64_t v = 0; v < (int64_t)n; ++v) dist[v] = kInf;
    dist[src] = 0;
}

Out top_down(const Csr& g, uint32_t* dist, const uint32_t* q, uint64_t size, uint64_t edges, uint32_t* q_next,
             uint32_t* deg_next, uint32_t level, int threads) {
    Out out;
    uint64_t tail = 0;
    const int nt = edges < kSerialEdges ? 1 : nthreads(threads);
#pragma omp parallel num_threads(nt)
    {
        static thread_local std::vector<uint32_t> lq, ld;
        lq.clear();
        ld.clear();
        uint64_t my_edges = 0, my_examined = 0;
#pragma omp for schedule(dynamic, 64) nowait
        for (int64_t i = 0; i < (int64_t)size; ++i) {
            const uint32_t v = q[i];
            my_examined += g.deg(v);
            for (const uint32_t* p = g.nbr_begin(v); p != g.nbr_end(v); ++p)
                if (claim(dist, *p, level + 1)) {
                    const uint32_t d = g.deg(*p);
                    lq.push_back(*p);
                    ld.push_back(d);
                    my_edges += d;
                }
        }
        const uint64_t at = __atomic_fetch_add(&tail, (uint64_t)lq.size(), __ATOMIC_RELAXED);
        if (!lq.empty()) {
            std::memcpy(q_next + at, lq.data(), lq.size() * 4);
            std::memcpy(deg_next + at, ld.data(), ld.size() * 4);
        }
        __atomic_fetch_add(&out.next_edges, my_edges, __ATOMIC_RELAXED);
        __atomic_fetch_add(&out.examined, my_examined, __ATOMIC_RELAXED);
    }
    out.found = tail;
    return out;
}

Out bottom_up(const Csr& g, uint32_t* dist, const uint64_t* front, uint64_t* next, uint32_t level, int threads) {
    const uint64_t n = g.n, words = (n + 63) / 64;
    uint64_t found = 0, examined = 0, next_edges = 0;
    // One thread owns whole words, so `next` is written without atomics.
#pragma omp parallel for num_threads(nthreads(threads)) schedule(dynamic, 32) \
    reduction(+ : found, examined, next_edges)
    for (int64_t w = 0; w < (int64_t)words; ++w) {
        uint64_t word = 0;
        const uint64_t hi = std::min<uint64_t>(n, (uint64_t)(w + 1) * 64);
        for (uint64_t v = (uint64_t)w * 64; v < hi; ++v) {
            if (dist[v] != kInf) continue;
            for (const uint32_t* p = g.nbr_begin(v); p != g.nbr_end(v); ++p) {
                ++examined;
                if ((front[*p >> 6] >> (*p & 63)) & 1ull) {
                    dist[v] = level + 1;
                    word |= 1ull << (v & 63);
                    ++found;
                    next_edges += g.deg(v);
                    break;
                }
            }
        }
        next[w] = word;
    }
    return {found, examined, next_edges};
}

void queue_to_bitmap(const uint32_t* q, uint64_t size, uint64_t* bm, uint64_t n, int threads) {
    const uint64_t words = (n + 63) / 64;
    const int nt = size < kSerialEdges ? 1 : nthreads(threads);
#pragma omp parallel num_threads(nt)
    {
#pragma omp for schedule(static)
        for (int64_t w = 0; w < (int64_t)words; ++w) bm[w] = 0;
#pragma omp for schedule(static)
        for (int64_t i = 0; i < (int64_t)size; ++i)
            __atomic_fetch_or(&bm[q[i] >> 6], 1ull << (q[i] & 63), __ATOMIC_RELAXED);
    }
}

uint64_t bitmap_to_queue(const Csr& g, const uint64_t* bm, uint32_t* q, uint32_t* deg, int threads) {
    const uint64_t words = (g.n + 63) / 64;
    const int nt = nthreads(threads);
    std::vector<uint64_t> start(nt + 1, 0);
    int team = 1;
    // Count per static word range, prefix, then write: increasing vertex order.
#pragma omp parallel num_threads(nt)
    {
        const int t = omp_get_thread_num(), T = omp_get_num_threads();
        const uint64_t lo = words * t / T, hi = words * (t + 1) / T;
        uint64_t c = 0;
        for (uint64_t w = lo; w < hi; ++w) c += (uint64_t)__builtin_popcountll(bm[w]);
        start[t + 1] = c;
#pragma omp barrier
#pragma omp single
        {
            team = T;
            for (int i = 0; i < T; ++i) start[i + 1] += start[i];
        }
        uint64_t at = start[t];
        for (uint64_t w = lo; w < hi; ++w)
            for (uint64_t bits = bm[w]; bits; bits &= bits - 1) {
                const uint32_t v = (uint32_t)(w * 64 + (uint64_t)__builtin_ctzll(bits));
                q[at] = v;
                deg[at] = g.deg(v);
                ++at;
            }
    }
    return start[team];
}

}  // namespace steps

}  // namespace ggpu
