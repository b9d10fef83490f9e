// Host model of the control protocol of the persistent single-source BFS
// loop (graph_analytics_applesilicon_amd/csrc/hip/bfs_loop.hip), test
// infrastructure only.
//
// Workgroups are coroutines that suspend before every access to memory that
// other workgroups share; a seeded scheduler interleaves them (with
// adversarial stalls). Memory has one cache per barrier group (the XCD a
// workgroup runs on under the round-robin workgroup -> XCD mapping):
//   plain load      hits the group's cache, else fills it from memory;
//   plain store     writes the group's cache (dirty);
//   store_through   writes memory and the group's copy;
//   coherent load   reads memory;
//   atomics         act on memory;
//   release fence   writes the group's dirty lines back: at once ("textbook")
//                   or queued and drained at random later steps ("late");
//   acquire fence   drops the group's clean lines.
// The decisions are the product's own code (fold_level / decide_action from
// bfs_engine.hpp, sb_ok and the command word from dl_protocol.hpp); the data
// plane (frontier / visited words) is abstracted: a level's counters come
// from the traversal's level profile (frontier size and edges per level),
// split over the workgroups.
//
// Two protocols:
//   old  every workgroup folds and decides itself from the per-block counters
//        (part slot (level + 1) & 1); workgroup 0 runs small top-down levels
//        alone and hands the state back with plain stores, a release fence
//        and a flag; the others wait for the flag, acquire and read it with
//        plain loads; levels end with the grid barrier (the previous kernel);
//   new  workgroup 0 alone decides and publishes commands (dl_protocol.hpp).
// Input (stdin): "n m" then "size edges" per level from level 0 until a size
// of 0; optional "script a0 a1 ..." line (scripted actions by level).
// Output: one JSON line per run.
#include <coroutine>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <exception>
#include <iostream>
#include <random>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "dl_protocol.hpp"
#include "grid_sync.hpp"

using namespace ggpu;

// ------------------------------------------------------------ coroutines
template <class T>
struct Co {
    struct promise_type {
        T value{};
        std::coroutine_handle<> cont;
        Co get_return_object() { return Co{std::coroutine_handle<promise_type>::from_promise(*this)}; }
        std::suspend_always initial_suspend() noexcept { return {}; }
        struct Final {
            bool await_ready() noexcept { return false; }
            std::coroutine_handle<> await_suspend(std::coroutine_handle<promise_type> h) noexcept {
                auto c = h.promise().cont;
                return c ? c : std::noop_coroutine();
            }
            void await_resume() noexcept {}
        };
        Final final_suspend() noexcept { return {}; }
        void return_value(T v) { value = v; }
        void unhandled_exception() { std::terminate(); }
    };
    std::coroutine_handle<promise_type> h;
    explicit Co(std::coroutine_handle<promise_type> x) : h(x) {}
    Co(Co&& o) noexcept : h(o.h) { o.h = nullptr; }
    Co(const Co&) = delete;
    ~Co() {
        if (h) h.destroy();
    }
    bool await_ready() { return false; }
    std::coroutine_handle<> await_suspend(std::coroutine_handle<> c) {
        h.promise().cont = c;
        return h;
    }
    T await_resume() { return h.promise().value; }
};

// ------------------------------------------------------------ memory
enum Mode { kTextbook = 0, kLate = 1 };

// addresses
constexpr uint64_t A_PART = 1ull << 40;  // + (slot * G + b) * 4 + field
constexpr uint64_t A_BAR = 2ull << 40;   // + word (grid_sync.hpp layout)
constexpr uint64_t A_ABORT = 3ull << 40;
constexpr uint64_t A_FLAG = 4ull << 40;  // old: single-workgroup episodes finished
constexpr uint64_t A_CMD = 5ull << 40;   // new: command word
constexpr uint64_t A_DS = 6ull << 40;    // old: hand-off state, one word per field

struct Line {
    uint64_t v;
    bool dirty;
};

struct Sim;

struct Yield {
    Sim* S;
    int b;
    bool await_ready() { return false; }
    void await_suspend(std::coroutine_handle<> h);
    void await_resume() {}
};

struct Block {
    std::coroutine_handle<> cur;
    bool done = false;
    uint64_t frozen_until = 0;
    std::vector<std::pair<uint32_t, uint32_t>> grid;  // grid levels run (level, action)
};

struct Sim {
    Mode mode;
    uint32_t G, groups;
    std::mt19937_64 rng;
    std::unordered_map<uint64_t, uint64_t> mem;
    std::vector<std::unordered_map<uint64_t, Line>> cache;
    std::deque<std::pair<uint32_t, uint64_t>> late;  // queued writebacks (group, address)
    std::vector<Block> blk;
    uint64_t step = 0, spin_limit = 100000;
    double p_stall = 0.0;       // after a barrier, probability that a workgroup stalls
    uint64_t stall_steps = 0;   // for this many scheduler steps
    // the traversal
    uint64_t n = 0, m = 0;
    std::vector<uint64_t> size, edges;
    std::vector<uint8_t> script;
    PolicyView pv{};

    uint32_t grp(uint32_t b) const { return b % groups; }
    uint64_t& M(uint64_t a) { return mem[a]; }
    uint64_t plain_load(uint32_t b, uint64_t a) {
        auto& c = cache[grp(b)];
        auto it = c.find(a);
        if (it != c.end()) return it->second.v;
        const uint64_t v = M(a);
        c[a] = Line{v, false};
        return v;
    }
    void plain_store(uint32_t b, uint64_t a, uint64_t v) { cache[grp(b)][a] = Line{v, true}; }
    void store_through(uint32_t b, uint64_t a, uint64_t v) {
        M(a) = v;
        auto& c = cache[grp(b)];
        auto it = c.find(a);
        if (it != c.end()) it->second = Line{v, false};
    }
    uint64_t cload(uint64_t a) { return M(a); }
    uint64_t atomic_add(uint64_t a, uint64_t v) {
        const uint64_t o = M(a);
        M(a) = o + v;
        return o;
    }
    void release(uint32_t b) {
        auto& c = cache[grp(b)];
        for (auto& kv : c)
            if (kv.second.dirty) {
                if (mode == kTextbook) {
                    M(kv.first) = kv.second.v;
                    kv.second.dirty = false;
                } else {
                    late.emplace_back(grp(b), kv.first);
                }
            }
    }
    void acquire(uint32_t b) {
        auto& c = cache[grp(b)];
        for (auto it = c.begin(); it != c.end();) {
            if (!it->second.dirty) it = c.erase(it);
            else ++it;
        }
    }
    void drain_one() {  // one queued or spontaneous writeback
        if (late.empty()) return;
        std::uniform_int_distribution<size_t> d(0, late.size() - 1);
        const size_t i = d(rng);
        const auto [g, a] = late[i];
        late.erase(late.begin() + (long)i);
        auto it = cache[g].find(a);
        if (it != cache[g].end() && it->second.dirty) {
            M(a) = it->second.v;
            it->second.dirty = false;
        }
    }
    Yield y(uint32_t b) { return Yield{this, (int)b}; }

    // level L's counters (found, next_edges) as workgroup b's share
    void share(uint32_t L, uint32_t b, uint64_t& found, uint64_t& ne) const {
        const uint64_t f = L + 1 < size.size() ? size[L + 1] : 0, e = L + 1 < edges.size() ? edges[L + 1] : 0;
        found = f * (b + 1) / G - f * b / G;
        ne = e * (b + 1) / G - e * b / G;
    }
    LoopState start() const {
        LoopState st{};
        st.size = 1;
        st.total = edges[0];
        st.scout = edges[0];
        st.to_check = m;
        st.unvisited_edges = m - edges[0];
        st.fmt = FMT_QUEUE;
        return st;
    }
};

void Yield::await_suspend(std::coroutine_handle<> h) { S->blk[b].cur = h; }

// ------------------------------------------------------------ shared pieces
static uint64_t part_addr(const Sim& S, uint32_t slot, uint32_t b, uint32_t f) {
    return A_PART + ((uint64_t)slot * S.G + b) * 4 + f;
}

// every block's counters of one slot, coherent loads (dl_totals)
static Co<int> totals(Sim& S, uint32_t b, uint32_t slot, LevelOut& r) {
    r = LevelOut{};
    for (uint32_t i = 0; i < S.G; ++i) {
        co_await S.y(b);
        r.found += S.cload(part_addr(S, slot, i, 0));
        co_await S.y(b);
        r.next_edges += S.cload(part_addr(S, slot, i, 1));
    }
    co_return 0;
}

static Co<int> write_part(Sim& S, uint32_t b, uint32_t slot, uint32_t level) {
    uint64_t f, e;
    S.share(level, b, f, e);
    co_await S.y(b);
    S.store_through(b, part_addr(S, slot, b, 0), f);
    co_await S.y(b);
    S.store_through(b, part_addr(S, slot, b, 1), e);
    co_return 0;
}

// grid_arrive: group counter, the last of a group bumps the top counter
static Co<int> arrive(Sim& S, uint32_t b) {
    const uint32_t g = b % kBarGroups, gsize = dl_group_size(S.G, kBarGroups, g);
    co_await S.y(b);
    if ((S.atomic_add(A_BAR + g * kBarStride, 1) + 1) % gsize == 0) {
        co_await S.y(b);
        S.atomic_add(A_BAR + kBarTop, 1);
    }
    co_return 0;
}

// grid_barrier (kAcquireL1 form of the previous kernel): 1 = passed, 0 = abort
static Co<int> barrier(Sim& S, uint32_t b) {
    co_await S.y(b);
    const uint64_t gen = S.cload(A_BAR + kBarGen);
    const uint32_t g = b % kBarGroups, gsize = dl_group_size(S.G, kBarGroups, g);
    bool released = false;
    co_await S.y(b);
    if ((S.atomic_add(A_BAR + g * kBarStride, 1) + 1) % gsize == 0) {
        co_await S.y(b);
        if ((S.atomic_add(A_BAR + kBarTop, 1) + 1) % S.groups == 0) {
            co_await S.y(b);
            S.atomic_add(A_BAR + kBarGen, 1);
            released = true;
        }
    }
    if (!released) {
        for (uint64_t spins = 0;;) {
            co_await S.y(b);
            if (S.cload(A_BAR + kBarGen) != gen) break;
            if (++spins > S.spin_limit) {
                S.M(A_ABORT) = 1;
                break;
            }
        }
    }
    if (S.p_stall > 0 && std::uniform_real_distribution<double>(0, 1)(S.rng) < S.p_stall)
        S.blk[b].frozen_until = S.step + S.stall_steps;
    co_await S.y(b);
    co_return S.cload(A_ABORT) == 0 ? 1 : 0;
}

// levels run by workgroup 0 alone, from the level profile (sb_levels):
// 1 = handed back with st decided for a level sb_ok rejects, 0 = traversal over
static Co<int> episode(Sim& S, LoopState& st, bool heartbeat, std::vector<std::pair<uint32_t, uint32_t>>& trace) {
    for (;;) {
        co_await S.y(0);
        if (heartbeat && (st.level & 63) == 0) S.store_through(0, A_CMD, dl_cmd(kCmdNone, st.level, st.action));
        trace.emplace_back(st.level, st.action | 0x100u);
        LevelOut r;
        r.found = st.level + 1 < S.size.size() ? S.size[st.level + 1] : 0;
        r.next_edges = st.level + 1 < S.edges.size() ? S.edges[st.level + 1] : 0;
        fold_level(st, r, S.pv, S.n);
        if (st.size == 0) co_return 0;
        decide_action(st, S.pv, S.n);
        if (!sb_ok(st)) co_return 1;
    }
}

// ------------------------------------------------------------ old protocol
// LoopState words handed off (plain stores / loads of D.ds->st, sb_more)
static const int kDsWords = (int)(sizeof(LoopState) / 8);

static Co<int> old_block(Sim& S, uint32_t b, std::vector<std::pair<uint32_t, uint32_t>>* trace0, uint32_t* final_level) {
    LoopState st;
    {  // kernel start: every workgroup reads the host-written state with plain loads
        uint64_t w[kDsWords];
        for (int i = 0; i < kDsWords; ++i) {
            co_await S.y(b);
            w[i] = S.plain_load(b, A_DS + i);
        }
        std::memcpy(&st, w, sizeof st);
    }
    bool started = false;
    uint32_t episodes = 0;
    for (;;) {
        if (started) {
            LevelOut r;
            co_await totals(S, b, (st.level + 1) & 1, r);
            fold_level(st, r, S.pv, S.n);
        }
        if (st.size == 0) break;
        decide_action(st, S.pv, S.n);
        started = true;
        if (sb_ok(st)) {
            ++episodes;
            if (b == 0) {
                co_await S.y(b);
                S.acquire(b);
                const int more = co_await episode(S, st, false, *trace0);
                uint64_t w[kDsWords];
                std::memcpy(w, &st, sizeof st);
                for (int i = 0; i < kDsWords; ++i) {
                    co_await S.y(b);
                    S.plain_store(b, A_DS + i, w[i]);
                }
                co_await S.y(b);
                S.plain_store(b, A_DS + kDsWords, (uint64_t)more);  // sb_more
                co_await S.y(b);
                S.release(b);
                co_await S.y(b);
                S.atomic_add(A_FLAG, 1);
                if (!more) break;
            } else {
                for (uint64_t spins = 0;;) {
                    co_await S.y(b);
                    if (S.cload(A_FLAG) >= episodes) break;
                    if (++spins > S.spin_limit || S.cload(A_ABORT)) {
                        S.M(A_ABORT) = 1;
                        co_return 0;
                    }
                }
                co_await S.y(b);
                S.acquire(b);
                co_await S.y(b);
                const uint64_t more = S.plain_load(b, A_DS + kDsWords);
                uint64_t w[kDsWords];
                for (int i = 0; i < kDsWords; ++i) {
                    co_await S.y(b);
                    w[i] = S.plain_load(b, A_DS + i);
                }
                std::memcpy(&st, w, sizeof st);
                if (!more) break;
            }
        }
        S.blk[b].grid.emplace_back(st.level, st.action);
        if (b == 0) trace0->emplace_back(st.level, st.action);
        co_await write_part(S, b, (st.level + 1) & 1, st.level);
        if (!co_await barrier(S, b)) co_return 0;
    }
    if (b == 0) *final_level = st.level;
    co_return 0;
}

// ------------------------------------------------------------ new protocol
static Co<int> new_block0(Sim& S, std::vector<std::pair<uint32_t, uint32_t>>* trace0, uint32_t* final_level) {
    LoopState st;
    {
        uint64_t w[kDsWords];
        for (int i = 0; i < kDsWords; ++i) {
            co_await S.y(0);
            w[i] = S.plain_load(0, A_DS + i);
        }
        std::memcpy(&st, w, sizeof st);
    }
    bool more = st.size != 0;
    if (more) decide_action(st, S.pv, S.n);
    for (uint32_t g = 0; more; ++g) {
        if (sb_ok(st)) {
            more = co_await episode(S, st, true, *trace0) != 0;
            if (!more) break;
        }
        co_await S.y(0);
        S.store_through(0, A_CMD, dl_cmd(kCmdGrid, st.level, st.action));
        S.blk[0].grid.emplace_back(st.level, st.action);
        trace0->emplace_back(st.level, st.action);
        co_await write_part(S, 0, dl_part_slot(g), st.level);
        co_await arrive(S, 0);
        for (uint64_t spins = 0;;) {
            co_await S.y(0);
            if (S.cload(A_BAR + kBarTop) >= (uint64_t)(g + 1) * S.groups) break;
            if (++spins > S.spin_limit || S.cload(A_ABORT)) {
                S.M(A_ABORT) = 1;
                S.store_through(0, A_CMD, dl_cmd(kCmdAbort, st.level, st.action));
                co_return 0;
            }
        }
        LevelOut r;
        co_await totals(S, 0, dl_part_slot(g), r);
        fold_level(st, r, S.pv, S.n);
        more = st.size != 0;
        if (more) decide_action(st, S.pv, S.n);
    }
    co_await S.y(0);
    S.store_through(0, A_CMD, dl_cmd(kCmdDone, st.level, 0));
    *final_level = st.level;
    co_return 0;
}

static Co<int> new_other(Sim& S, uint32_t b) {
    uint32_t level1 = 0;
    for (uint32_t g = 0;; ++g) {
        uint64_t c = 0, seen = ~0ull;
        for (uint64_t spins = 0;;) {
            co_await S.y(b);
            c = S.cload(A_CMD);
            const uint32_t k = dl_cmd_kind(c);
            if (k == kCmdDone || k == kCmdAbort) co_return 0;
            if (k == kCmdGrid && dl_cmd_level1(c) > level1) break;
            if (c != seen) {
                seen = c;
                spins = 0;
            } else if (++spins > S.spin_limit || S.cload(A_ABORT)) {
                S.M(A_ABORT) = 1;
                co_return 0;
            }
        }
        level1 = dl_cmd_level1(c);
        S.blk[b].grid.emplace_back(level1 - 1, dl_cmd_action(c));
        co_await write_part(S, b, dl_part_slot(g), level1 - 1);
        co_await arrive(S, b);
    }
}

// ------------------------------------------------------------ driver
int main(int argc, char** argv) {
    std::string proto = "new";
    Mode mode = kTextbook;
    uint64_t seed = 1, runs = 1;
    uint32_t G = 256;
    double p_stall = 0.0, p_drain = 0.02;
    uint64_t stall_steps = 0, spin = 100000;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() { return std::string(i + 1 < argc ? argv[++i] : ""); };
        if (a == "--protocol") proto = next();
        else if (a == "--mode") mode = next() == "late" ? kLate : kTextbook;
        else if (a == "--seed") seed = std::stoull(next());
        else if (a == "--runs") runs = std::stoull(next());
        else if (a == "--grid") G = (uint32_t)std::stoul(next());
        else if (a == "--stall") p_stall = std::stod(next());
        else if (a == "--stall-steps") stall_steps = std::stoull(next());
        else if (a == "--drain") p_drain = std::stod(next());
        else if (a == "--spin") spin = std::stoull(next());
        else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    if (proto != "old" && proto != "new") return 2;
    uint64_t n = 0, m = 0;
    std::vector<uint64_t> size, edges;
    std::vector<uint8_t> script;
    {
        std::string line;
        std::getline(std::cin, line);
        std::istringstream(line) >> n >> m;
        while (std::getline(std::cin, line)) {
            std::istringstream is(line);
            std::string w;
            is >> w;
            if (w == "script") {
                int x;
                while (is >> x) script.push_back((uint8_t)x);
                continue;
            }
            if (w.empty()) continue;
            size.push_back(std::stoull(w));
            uint64_t e = 0;
            is >> e;
            edges.push_back(e);
        }
    }
    if (size.empty() || size[0] != 1) {
        std::fprintf(stderr, "level profile must start with a frontier of 1\n");
        return 2;
    }
    // the sequential decisions (the host loop's): level, action, run alone (0x100)
    std::vector<std::pair<uint32_t, uint32_t>> ref;
    for (uint64_t r = 0; r < runs; ++r) {
        Sim S;
        S.mode = mode;
        S.G = G;
        S.groups = dl_groups(G, kBarGroups);
        S.rng.seed(seed + r);
        S.cache.resize(S.groups);
        S.blk.resize(G);
        S.p_stall = p_stall;
        S.stall_steps = stall_steps;
        S.spin_limit = spin;
        S.n = n;
        S.m = m;
        S.size = size;
        S.edges = edges;
        S.script = script;
        S.pv.mode = script.empty() ? GGPU_POLICY_HEURISTIC : GGPU_POLICY_SCRIPTED;
        S.pv.variant = GGPU_BFS_DO;
        S.pv.alpha = 15;
        S.pv.beta = 18;
        S.pv.actions = S.script.empty() ? nullptr : S.script.data();
        S.pv.n_actions = (uint32_t)S.script.size();
        if (ref.empty()) {
            LoopState st = S.start();
            decide_action(st, S.pv, n);
            for (;;) {
                ref.emplace_back(st.level, st.action | (sb_ok(st) ? 0x100u : 0u));
                LevelOut lo;
                lo.found = st.level + 1 < size.size() ? size[st.level + 1] : 0;
                lo.next_edges = st.level + 1 < edges.size() ? edges[st.level + 1] : 0;
                fold_level(st, lo, S.pv, n);
                if (st.size == 0) break;
                decide_action(st, S.pv, n);
            }
        }
        {  // the host writes the initial state before the launch
            LoopState st = S.start();
            uint64_t w[kDsWords];
            std::memcpy(w, &st, sizeof st);
            for (int i = 0; i < kDsWords; ++i) S.M(A_DS + i) = w[i];
        }
        std::vector<std::pair<uint32_t, uint32_t>> trace0;
        uint32_t final_level = 0;
        std::vector<Co<int>> tasks;
        tasks.reserve(G);
        for (uint32_t b = 0; b < G; ++b) {
            if (proto == "old") tasks.push_back(old_block(S, b, &trace0, &final_level));
            else tasks.push_back(b == 0 ? new_block0(S, &trace0, &final_level) : new_other(S, b));
            S.blk[b].cur = tasks.back().h;
        }
        std::uniform_int_distribution<uint32_t> pick(0, G - 1);
        std::uniform_real_distribution<double> u(0, 1);
        uint32_t live = G;
        bool hang = false;
        const uint64_t cap = 4000000000ull;
        while (live) {
            if (++S.step > cap) {
                hang = true;
                break;
            }
            if (S.mode == kLate && !S.late.empty() && u(S.rng) < p_drain) S.drain_one();
            uint32_t b = pick(S.rng);
            if (S.blk[b].done || S.blk[b].frozen_until > S.step) {
                // find a runnable block (frozen blocks wait their turn)
                uint32_t k = 0;
                for (; k < G; ++k) {
                    const uint32_t c = (b + k) % G;
                    if (!S.blk[c].done && S.blk[c].frozen_until <= S.step) break;
                }
                if (k == G) continue;  // everyone left is frozen: time passes
                b = (b + k) % G;
            }
            S.blk[b].cur.resume();
            if (tasks[b].h.done()) {
                S.blk[b].done = true;
                --live;
            }
        }
        // outcome
        const bool aborted = S.M(A_ABORT) != 0;
        bool same_grid = true;
        for (uint32_t b = 1; b < G; ++b)
            if (S.blk[b].grid != S.blk[0].grid) same_grid = false;
        const bool trace_ok = trace0 == ref && final_level == ref.back().first + 1;
        const char* outcome = hang ? "hang" : aborted ? "timeout" : (!same_grid ? "diverged" : (!trace_ok ? "wrong" : "ok"));
        std::printf("{\"protocol\": \"%s\", \"mode\": \"%s\", \"seed\": %llu, \"grid\": %u, \"outcome\": \"%s\", "
                    "\"group_arrivals\": [",
                    proto.c_str(), mode == kLate ? "late" : "textbook", (unsigned long long)(seed + r), G, outcome);
        for (uint32_t g = 0; g < kBarGroups; ++g)
            std::printf("%s%llu", g ? ", " : "", (unsigned long long)S.M(A_BAR + g * kBarStride));
        std::printf("], \"top\": %llu, \"generation\": %llu, \"episodes\": %llu, \"level\": %u, \"levels\": %zu, "
                    "\"grid_levels\": %zu, \"exited_groups\": [",
                    (unsigned long long)S.M(A_BAR + kBarTop), (unsigned long long)S.M(A_BAR + kBarGen),
                    (unsigned long long)S.M(A_FLAG), final_level, ref.size(), S.blk[0].grid.size());
        // groups whose every workgroup ran fewer grid levels than the most any ran
        size_t most = 0;
        for (auto& B : S.blk) most = std::max(most, B.grid.size());
        bool first = true;
        for (uint32_t g = 0; g < S.groups; ++g) {
            bool all_short = true;
            for (uint32_t b = g; b < G; b += S.groups) all_short &= S.blk[b].grid.size() < most;
            if (all_short) {
                std::printf("%s%u", first ? "" : ", ", g);
                first = false;
            }
        }
        std::printf("], \"steps\": %llu}\n", (unsigned long long)S.step);
        std::fflush(stdout);
    }
    return 0;
}
