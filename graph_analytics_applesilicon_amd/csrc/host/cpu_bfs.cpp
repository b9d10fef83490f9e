// CPU BFS variants and the hybrid BFS's CPU levels.
//   bfs_serial  src/cpu/bfs_cpu.cpp:27-45   FIFO BFS (gbench --validate's truth)
//   bfs_omp     src/cpu/bfs_cpu.cpp:207-223 every level top-down (cpu_omp)
//   bfs_do      src/cpu/bfs_cpu.cpp:227-283 direction-optimizing (cpu_do)
//   steps::*    src/cpu/bfs_steps.cpp:27-143, the same contract on the GPU's
//               buffer layout (queue + degrees top-down, u64 bitmaps bottom-up)
//
// Design:
//  * a frontier is a bitmap of u64 words with one summary bit per word: a
//    frontier is walked, cleared and collected through its summary, touching
//    only its non-zero words;
//  * cpu_omp / cpu_do discover a vertex with the fetch_or on the visited
//    bitmap that sets its bit (exactly once; dist gets a plain store), and a
//    level sets the next frontier's bits; no per-thread queues;
//  * a queue made by several threads is collected from a bitmap in
//    increasing vertex order, and a one-thread top-down step appends in
//    discovery order, so no level's output depends on thread timing;
//  * work is dealt in units (words of 64 vertices, or queue entries) by one
//    atomic cursor: a static split, chunks of 64 units, or guided halving
//    (--schedule static|dynamic|guided);
//  * cpu_do's direction per level comes from the GPU engine's own state
//    machine (fold_level / decide_action, level_policy.hpp: the GAP rule),
//    so its level sequence is the one every GPU loop follows.
#include "graph.hpp"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cstring>

#include "level_policy.hpp"

namespace ggpu {

namespace {

int team(int t) { return t > 0 ? t : omp_get_max_threads(); }

double now_ms() { return omp_get_wtime() * 1e3; }

struct Counts {
    uint64_t found = 0, examined = 0, next_edges = 0;
    void add(const Counts& o) {
        found += o.found;
        examined += o.examined;
        next_edges += o.next_edges;
    }
};

// Runs body(first, end, counts) over [0, units) on nt threads under policy
// `pol` (Dynamic64: chunks of `chunk` units); returns the summed counts.
template <class Body>
Counts deal(int nt, Schedule pol, uint64_t units, Body body, uint64_t chunk_units = 64) {
    Counts total;
    if (units == 0) return total;
    if (nt <= 1 || units < 2) {
        body(0, units, total);
        return total;
    }
    std::atomic<uint64_t> cursor{0};
#pragma omp parallel num_threads(nt)
    {
        Counts mine;
        const uint64_t t = (uint64_t)omp_get_thread_num(), T = (uint64_t)omp_get_num_threads();
        if (pol == Schedule::Static) {
            body(units * t / T, units * (t + 1) / T, mine);
        } else {
            for (;;) {
                uint64_t chunk = chunk_units;
                if (pol == Schedule::Guided) {
                    const uint64_t seen = cursor.load(std::memory_order_relaxed);
                    chunk = std::max<uint64_t>(chunk_units, (units - std::min(units, seen)) / (2 * T));
                }
                const uint64_t b = cursor.fetch_add(chunk, std::memory_order_relaxed);
                if (b >= units) break;
                body(b, std::min(units, b + chunk), mine);
            }
        }
#pragma omp critical(ggpu_cpu_bfs_counts)
        total.add(mine);
    }
    return total;
}

inline bool test_bit(const uint64_t* w, uint32_t v) { return (w[v >> 6] >> (v & 63)) & 1ull; }

// A vertex set as words plus a summary bit per word (bit i%64 of sum[i/64]:
// word i is not zero), so walking, clearing and collecting a sparse set touch
// only its words.
struct Frontier {
    uint64_t n = 0, nw = 0;
    std::vector<uint64_t> w, sum;
    mutable std::vector<uint64_t> at;    // collect's prefix of summary words
    explicit Frontier(uint64_t n_ = 0) { resize(n_); }
    void resize(uint64_t n_) {
        n = n_;
        nw = (n + 63) / 64;
        w.assign(nw, 0);
        sum.assign((nw + 63) / 64, 0);
    }
    // inserts; Shared: other threads insert concurrently
    template <bool Shared>
    void insert(uint32_t v) {
        uint64_t old;
        if (Shared) {
            old = __atomic_fetch_or(&w[v >> 6], 1ull << (v & 63), __ATOMIC_RELAXED);
        } else {
            old = w[v >> 6];
            w[v >> 6] = old | 1ull << (v & 63);
        }
        if (old == 0) {
            if (Shared) __atomic_fetch_or(&sum[v >> 12], 1ull << ((v >> 6) & 63), __ATOMIC_RELAXED);
            else sum[v >> 12] |= 1ull << ((v >> 6) & 63);
        }
    }
    // a whole word written by its only writer
    void put_word(uint64_t i, uint64_t bits) {
        w[i] = bits;
        if (bits) __atomic_fetch_or(&sum[i >> 6], 1ull << (i & 63), __ATOMIC_RELAXED);
    }
    // words that are not zero (counted when needed: a shared counter updated by
    // every insert costs the parallel levels a contended atomic per new word)
    uint64_t words_set() const {
        uint64_t c = 0;
        for (uint64_t x : sum) c += (uint64_t)__builtin_popcountll(x);
        return c;
    }
    // sparse sets are cleared and collected by one thread
    int team_for(int nt) const { return words_set() <= 4096 ? 1 : nt; }
    void clear(int nt) {
        deal(team_for(nt), Schedule::Dynamic64, sum.size(), [&](uint64_t b, uint64_t e, Counts&) {
            for (uint64_t s = b; s < e; ++s) {
                for (uint64_t bits = sum[s]; bits; bits &= bits - 1) w[s * 64 + (uint64_t)__builtin_ctzll(bits)] = 0;
                sum[s] = 0;
            }
        }, 8);
    }
    // the members in increasing order (q) with their degrees (deg, optional)
    uint64_t collect(const Csr& g, uint32_t* q, uint32_t* deg, int nt) const {
        nt = team_for(nt);
        const uint64_t ns = sum.size();
        if (nt == 1) {  // one pass
            uint64_t k = 0;
            for (uint64_t s = 0; s < ns; ++s)
                for (uint64_t bits = sum[s]; bits; bits &= bits - 1) {
                    const uint64_t i = s * 64 + (uint64_t)__builtin_ctzll(bits);
                    for (uint64_t m = w[i]; m; m &= m - 1) {
                        const uint32_t v = (uint32_t)(i * 64 + (uint64_t)__builtin_ctzll(m));
                        q[k] = v;
                        if (deg) deg[k] = g.deg(v);
                        ++k;
                    }
                }
            return k;
        }
        at.assign(ns + 1, 0);
        deal(nt, Schedule::Dynamic64, ns, [&](uint64_t b, uint64_t e, Counts&) {
            for (uint64_t s = b; s < e; ++s) {
                uint64_t c = 0;
                for (uint64_t bits = sum[s]; bits; bits &= bits - 1)
                    c += (uint64_t)__builtin_popcountll(w[s * 64 + (uint64_t)__builtin_ctzll(bits)]);
                at[s + 1] = c;
            }
        }, 8);
        for (uint64_t s = 0; s < ns; ++s) at[s + 1] += at[s];
        deal(nt, Schedule::Dynamic64, ns, [&](uint64_t b, uint64_t e, Counts&) {
            for (uint64_t s = b; s < e; ++s) {
                uint64_t k = at[s];
                for (uint64_t bits = sum[s]; bits; bits &= bits - 1) {
                    const uint64_t i = s * 64 + (uint64_t)__builtin_ctzll(bits);
                    for (uint64_t m = w[i]; m; m &= m - 1) {
                        const uint32_t v = (uint32_t)(i * 64 + (uint64_t)__builtin_ctzll(m));
                        q[k] = v;
                        if (deg) deg[k] = g.deg(v);
                        ++k;
                    }
                }
            }
        }, 8);
        return at[ns];
    }
};

// Top-down from the set `cur` into `next`, dealt by summary words (4096
// vertices each, 8 per chunk); `vis` decides discovery. examined = the
// frontier's degrees. `cur` is left empty: each word is cleared by the thread
// that walked it.
template <bool Shared>
Counts td_walk(const Csr& g, Frontier& cur, Frontier& next, uint64_t* vis, uint32_t* dist, uint32_t depth, int nt,
               Schedule pol) {
    return deal(
        nt, pol, cur.sum.size(),
        [&](uint64_t b, uint64_t e, Counts& c) {
            for (uint64_t s = b; s < e; ++s) {
                const uint64_t words = cur.sum[s];
                cur.sum[s] = 0;
                for (uint64_t wb = words; wb; wb &= wb - 1) {
                    const uint64_t i = s * 64 + (uint64_t)__builtin_ctzll(wb);
                    const uint64_t word = cur.w[i];
                    cur.w[i] = 0;
                    for (uint64_t bits = word; bits; bits &= bits - 1) {
                        const uint32_t v = (uint32_t)(i * 64 + (uint64_t)__builtin_ctzll(bits));
                        c.examined += g.deg(v);
                        for (const uint32_t* p = g.nbr_begin(v); p != g.nbr_end(v); ++p) {
                            const uint32_t u = *p;
                            const uint64_t bit = 1ull << (u & 63);
                            if (Shared) {
                                if (__atomic_load_n(&vis[u >> 6], __ATOMIC_RELAXED) & bit) continue;
                                if (__atomic_fetch_or(&vis[u >> 6], bit, __ATOMIC_RELAXED) & bit) continue;
                            } else {
                                if (vis[u >> 6] & bit) continue;
                                vis[u >> 6] |= bit;
                            }
                            dist[u] = depth;
                            next.insert<Shared>(u);
                            c.found += 1;
                            c.next_edges += g.deg(u);
                        }
                    }
                }
            }
        },
        8);
}

Counts td_from_set(const Csr& g, Frontier& cur, Frontier& next, uint64_t* vis, uint32_t* dist, uint32_t depth,
                   int nt, Schedule pol) {
    return nt > 1 ? td_walk<true>(g, cur, next, vis, dist, depth, nt, pol)
                  : td_walk<false>(g, cur, next, vis, dist, depth, 1, pol);
}

// Bottom-up: each unvisited vertex takes the first neighbour in `cur`; the
// thread handling word i of the vertex range owns vis[i] and next's word i.
Counts bu_to_set(const Csr& g, const Frontier& cur, Frontier& next, uint64_t* vis, uint32_t* dist, uint32_t depth,
                 int nt, Schedule pol) {
    const uint64_t n = g.n, nw = (n + 63) / 64;
    return deal(nt, pol, nw, [&](uint64_t b, uint64_t e, Counts& c) {
        for (uint64_t i = b; i < e; ++i) {
            uint64_t open = ~vis[i];
            if (i == nw - 1 && (n & 63)) open &= (1ull << (n & 63)) - 1;
            uint64_t got = 0;
            for (; open; open &= open - 1) {
                const uint32_t v = (uint32_t)(i * 64 + (uint64_t)__builtin_ctzll(open));
                const uint32_t* p = g.nbr_begin(v);
                const uint32_t* end = g.nbr_end(v);
                for (; p != end && !test_bit(cur.w.data(), *p); ++p) {
                }
                c.examined += (uint64_t)(p - g.nbr_begin(v)) + (p != end ? 1 : 0);
                if (p != end) {
                    got |= 1ull << (v & 63);
                    dist[v] = depth;
                    c.next_edges += g.deg(v);
                }
            }
            if (got) {
                vis[i] |= got;
                next.put_word(i, got);
                c.found += (uint64_t)__builtin_popcountll(got);
            }
        }
    });
}

ggpu_level make_level(uint32_t level, uint64_t frontier, uint64_t edges, bool bottom_up, double ms) {
    ggpu_level L{};
    L.level = level;
    L.action = bottom_up ? GGPU_ACTION_BU : GGPU_ACTION_TD_EDGE;
    L.frontier = frontier;
    L.edges = edges;
    L.frontier_edges = bottom_up ? 0 : edges;
    L.direction = bottom_up ? 'B' : 'T';
    L.device = 'C';
    L.time_ms = ms;
    return L;
}

void fill_inf(uint32_t* dist, uint64_t n, int nt) {
    deal(nt, Schedule::Static, (n + 4095) / 4096, [&](uint64_t b, uint64_t e, Counts&) {
        std::fill(dist + b * 4096, dist + std::min<uint64_t>(n, e * 4096), kInf);
    });
}

// Level loop of cpu_omp (all top-down) and cpu_do (the GAP state machine).
void level_loop(const Csr& g, uint32_t src, uint32_t* dist, const CpuBfsOpts& o, std::vector<ggpu_level>* trace,
                bool direction_optimizing) {
    const int nt = team(o.threads);
    const uint64_t n = g.n;
    fill_inf(dist, n, nt);
    std::vector<uint64_t> vis((n + 63) / 64, 0);
    Frontier cur(n), next(n);
    dist[src] = 0;
    vis[src >> 6] |= 1ull << (src & 63);
    cur.insert<false>(src);
    LoopState st{};
    st.size = 1;
    st.total = g.deg(src);
    st.scout = st.total;
    st.to_check = g.m;
    st.unvisited_edges = g.m - st.total;
    PolicyView pv{};
    pv.mode = GGPU_POLICY_HEURISTIC;
    pv.variant = GGPU_BFS_DO;
    pv.alpha = o.alpha;
    pv.beta = o.beta;
    // cpu_do deals chunks of 64 words (the reference's dynamic,64); cpu_omp follows --schedule
    const Schedule pol = direction_optimizing ? Schedule::Dynamic64 : o.schedule;
    for (;;) {
        if (direction_optimizing) decide_action(st, pv, n);
        else st.action = GGPU_ACTION_TD_EDGE;
        const bool bu = st.action == GGPU_ACTION_BU;
        const double t0 = trace ? now_ms() : 0.0;
        // a top-down level of few edges runs on one thread (no parallel region)
        const int lnt = !bu && st.total < steps::kSerialEdges ? 1 : nt;
        const Counts c = bu ? bu_to_set(g, cur, next, vis.data(), dist, st.level + 1, nt, pol)
                            : td_from_set(g, cur, next, vis.data(), dist, st.level + 1, lnt, pol);
        if (bu) cur.clear(nt);  // (a top-down level has emptied cur itself)
        std::swap(cur, next);
        if (trace) trace->push_back(make_level(st.level, st.size, c.examined, bu, now_ms() - t0));
        LevelOut r;
        r.found = c.found;
        r.examined = c.examined;
        r.next_edges = c.next_edges;
        fold_level(st, r, pv, n);
        if (st.size == 0) break;
    }
}

// The calling thread's scratch set of the hybrid's top-down level.
Frontier& scratch_set(uint64_t n) {
    static thread_local Frontier f;
    if (f.n != n) f.resize(n);
    return f;
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
    level_loop(g, src, dist, o, trace, false);
}

void bfs_do(const Csr& g, uint32_t src, uint32_t* dist, const CpuBfsOpts& o, std::vector<ggpu_level>* trace) {
    level_loop(g, src, dist, o, trace, true);
}

namespace steps {

void init_dist(uint32_t* dist, uint64_t n, uint32_t src, int threads) {
    fill_inf(dist, n, n > (1u << 16) ? team(threads) : 1);
    dist[src] = 0;
}

// The hybrid mirror keeps no visited bitmap: discovery is the first write of
// dist from INF. A small level (few edges) runs on one thread and appends each
// discovery to the next queue as it is made (that order depends only on the
// input queue); a large one claims with a CAS, puts discoveries in the thread's
// scratch set and collects them in vertex order.
Out top_down(const Csr& g, uint32_t* dist, const uint32_t* q, uint64_t size, uint64_t edges, uint32_t* q_next,
             uint32_t* deg_next, uint32_t level, int threads) {
    const int nt = edges < kSerialEdges ? 1 : team(threads);
    const uint32_t depth = level + 1;
    Out out;
    if (nt == 1) {
        uint64_t found = 0;
        const Counts c = deal(1, Schedule::Static, size, [&](uint64_t b, uint64_t e, Counts& k) {
            for (uint64_t i = b; i < e; ++i) {
                const uint32_t v = q[i];
                k.examined += g.deg(v);
                for (const uint32_t* p = g.nbr_begin(v); p != g.nbr_end(v); ++p) {
                    if (dist[*p] != kInf) continue;
                    dist[*p] = depth;
                    const uint32_t dg = g.deg(*p);
                    q_next[found] = *p;
                    deg_next[found] = dg;
                    ++found;
                    k.next_edges += dg;
                }
            }
        });
        out.found = found;
        out.examined = c.examined;
        out.next_edges = c.next_edges;
        return out;
    }
    Frontier& S = scratch_set(g.n);
    const Counts c = deal(nt, Schedule::Dynamic64, size, [&](uint64_t b, uint64_t e, Counts& k) {
        for (uint64_t i = b; i < e; ++i) {
            const uint32_t v = q[i];
            k.examined += g.deg(v);
            for (const uint32_t* p = g.nbr_begin(v); p != g.nbr_end(v); ++p) {
                uint32_t seen = __atomic_load_n(&dist[*p], __ATOMIC_RELAXED);
                if (seen != kInf ||
                    !__atomic_compare_exchange_n(&dist[*p], &seen, depth, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED))
                    continue;
                S.insert<true>(*p);
                k.next_edges += g.deg(*p);
            }
        }
    });
    out.found = S.collect(g, q_next, deg_next, nt);
    out.examined = c.examined;
    out.next_edges = c.next_edges;
    S.clear(nt);
    return out;
}

// Unvisited = dist INF. Every word of `next` is written (the GPU layout reads it whole).
Out bottom_up(const Csr& g, uint32_t* dist, const uint64_t* front, uint64_t* next, uint32_t level, int threads) {
    const uint64_t n = g.n, nw = (n + 63) / 64;
    const uint32_t depth = level + 1;
    const Counts c = deal(team(threads), Schedule::Dynamic64, nw, [&](uint64_t b, uint64_t e, Counts& k) {
        for (uint64_t i = b; i < e; ++i) {
            const uint64_t lo = i * 64, hi = std::min<uint64_t>(n, lo + 64);
            uint64_t open = 0;
            for (uint64_t v = lo; v < hi; ++v) open |= (uint64_t)(dist[v] == kInf) << (v - lo);
            uint64_t got = 0;
            for (; open; open &= open - 1) {
                const uint32_t v = (uint32_t)(lo + (uint64_t)__builtin_ctzll(open));
                const uint32_t* p = g.nbr_begin(v);
                const uint32_t* end = g.nbr_end(v);
                for (; p != end && !test_bit(front, *p); ++p) {
                }
                k.examined += (uint64_t)(p - g.nbr_begin(v)) + (p != end ? 1 : 0);
                if (p != end) {
                    got |= 1ull << (v - lo);
                    dist[v] = depth;
                    k.next_edges += g.deg(v);
                }
            }
            next[i] = got;
            k.found += (uint64_t)__builtin_popcountll(got);
        }
    });
    Out out;
    out.found = c.found;
    out.examined = c.examined;
    out.next_edges = c.next_edges;
    return out;
}

void queue_to_bitmap(const uint32_t* q, uint64_t size, uint64_t* bm, uint64_t n, int threads) {
    const int nt = size < kSerialEdges ? 1 : team(threads);
    std::memset(bm, 0, ((n + 63) / 64) * 8);
    deal(nt, Schedule::Dynamic64, size, [&](uint64_t b, uint64_t e, Counts&) {
        for (uint64_t i = b; i < e; ++i) __atomic_fetch_or(&bm[q[i] >> 6], 1ull << (q[i] & 63), __ATOMIC_RELAXED);
    });
}

// Increasing vertex order: popcounts of fixed 4096-word spans, their prefix, then the writes.
uint64_t bitmap_to_queue(const Csr& g, const uint64_t* bm, uint32_t* q, uint32_t* deg, int threads) {
    const uint64_t nw = (g.n + 63) / 64, spans = (nw + 4095) / 4096;
    const int nt = team(threads);
    std::vector<uint64_t> at(spans + 1, 0);
    deal(nt, Schedule::Dynamic64, spans, [&](uint64_t b, uint64_t e, Counts&) {
        for (uint64_t s = b; s < e; ++s) {
            uint64_t c = 0;
            for (uint64_t i = s * 4096; i < std::min(nw, s * 4096 + 4096); ++i) c += (uint64_t)__builtin_popcountll(bm[i]);
            at[s + 1] = c;
        }
    });
    for (uint64_t s = 0; s < spans; ++s) at[s + 1] += at[s];
    deal(nt, Schedule::Dynamic64, spans, [&](uint64_t b, uint64_t e, Counts&) {
        for (uint64_t s = b; s < e; ++s) {
            uint64_t k = at[s];
            for (uint64_t i = s * 4096; i < std::min(nw, s * 4096 + 4096); ++i)
                for (uint64_t bits = bm[i]; bits; bits &= bits - 1) {
                    const uint32_t v = (uint32_t)(i * 64 + (uint64_t)__builtin_ctzll(bits));
                    q[k] = v;
                    deg[k] = g.deg(v);
                    ++k;
                }
        }
    });
    return at[spans];
}

}  // namespace steps

}  // namespace ggpu
