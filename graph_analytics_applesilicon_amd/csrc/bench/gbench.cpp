// gbench: benchmark + validation CLI over the two C-ABIs (ggpu_graph.h,
// ggpu_hip.h). It keeps the reference harness's command line, methodology,
// console summary and CSV schema (src/bench/main.cpp:1-27, :205-420,
// :422-682), so scripts written for it run unchanged:
//
//   gbench --graph FILE.bin|synth:NAME --algo bfs --variant V
//          [--threads N] [--schedule static|dynamic,64|guided] [--alpha 15 --beta 18]
//          [--tg 256] [--small-max 16 --large-min 256] [--ept 2] [--T 100000|inf]
//          [--sources 64 | --source S] [--seed 1] [--any-component] [--warmup 2] [--trials 5]
//          [--source-warmup 1] [--validate] [--trace trace.csv] [--csv out.csv] [--dump-dist out.u32]
//          [--loop host|graph|persistent] [--policy heuristic|learned] [--policy-file W.f32]
//   gbench --graph X --algo pr --variant V [--iters 20 | --tol 1e-6 [--iters 1000]] [--tol-rel R]
//          [--tg 256] [--small-max 16 --large-min 256] [--inplace] [--ref-tol T] [--gs-pairs 2]
//          [--threads N] [--warmup 2] [--trials 5] [--validate] [--csv out.csv] [--dump-ranks out.f64]
//   gbench --graph X --algo msbfs [--batch 64] [--sources 64] [--policy heuristic|learned]
//          [--batch-order auto|degree|given]
//          [--policy-file W.f32] [--warmup 2] [--trials 5] [--validate] [--csv out.csv]
//   every algorithm: [--ordering original|random|degree|rcm] [--order-seed 1] [--locality-csv loc.csv]
//          (vertices relabeled before the run, untimed; results mapped back for validation)
//   gbench --graph X --save-graph out.bin
//
// BFS variants: cpu_serial cpu_omp cpu_do (host library) and gpu_vertex
// gpu_naive gpu_binned gpu_edge gpu_do hybrid (HIP backend). PageRank:
// cpu_serial cpu_omp and gpu_push gpu_pull_scalar gpu_pull_vector
// gpu_pull_adaptive gpu_pull_gs. msbfs (no
// reference counterpart) runs the sources as batches of --batch in one launch
// each. Exit codes as in the reference: 0 ok, 1 failure / validation
// mismatch, 2 usage (unknown variant or algorithm, bad source).
#include <omp.h>
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cli.hpp"
#include "ggpu_graph.h"
#include "ggpu_hip.h"

namespace {

// ------------------------------------------------------------ small helpers
struct Failure : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class Args {
   public:
    Args(int argc, char** argv) {
        for (int i = 1; i < argc; ++i) {
            std::string a = argv[i];
            if (a.rfind("--", 0) != 0) throw Failure("unexpected argument '" + a + "'");
            a = a.substr(2);
            if (i + 1 < argc && std::strncmp(argv[i + 1], "--", 2) != 0) kv_[a] = argv[++i];
            else kv_[a] = "";
        }
    }
    bool has(const std::string& k) const { return kv_.count(k) != 0; }
    std::string get(const std::string& k, const std::string& d) const {
        auto it = kv_.find(k);
        return it == kv_.end() ? d : it->second;
    }
    long long num(const std::string& k, long long d) const { return has(k) ? std::stoll(get(k, "")) : d; }
    double real(const std::string& k, double d) const { return has(k) ? std::stod(get(k, "")) : d; }

   private:
    std::map<std::string, std::string> kv_;
};

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

double median(std::vector<double> v) {
    if (v.empty()) return 0;
    std::sort(v.begin(), v.end());
    const size_t k = v.size();
    return k % 2 ? v[k / 2] : 0.5 * (v[k / 2 - 1] + v[k / 2]);
}

bool file_exists(const std::string& p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0;
}

std::string trim(std::string s) {
    while (!s.empty() && (s.back() == '\n' || s.back() == '\r' || s.back() == ' ')) s.pop_back();
    return s;
}

std::string run_cmd(const char* cmd) {
    std::string out;
    if (FILE* f = popen(cmd, "r")) {
        char buf[256];
        while (std::fgets(buf, sizeof buf, f)) out += buf;
        pclose(f);
    }
    return trim(out);
}

std::string timestamp() {
    std::time_t t = std::time(nullptr);
    char b[32];
    std::strftime(b, sizeof b, "%Y-%m-%dT%H:%M:%S", std::localtime(&t));
    return b;
}

double load1() {
    std::ifstream f("/proc/loadavg");
    double x = 0;
    f >> x;
    return x;
}

std::string cpu_name() {
    std::ifstream f("/proc/cpuinfo");
    std::string line;
    while (std::getline(f, line))
        if (line.rfind("model name", 0) == 0) return trim(line.substr(line.find(':') + 2));
    return "cpu";
}

std::string no_commas(std::string s) {
    std::replace(s.begin(), s.end(), ',', ' ');
    return s;
}

void host_check(int rc, const char* what) {
    if (rc != GGPU_OK) throw Failure(std::string(what) + ": " + ggpu_host_last_error());
}

// Owning handles over the C-ABI objects.
struct Graph {
    ggpu_graph* g = nullptr;
    Graph() = default;
    explicit Graph(ggpu_graph* p) : g(p) {}
    Graph(const Graph&) = delete;
    Graph& operator=(const Graph&) = delete;
    Graph(Graph&& o) noexcept : g(o.g) { o.g = nullptr; }
    Graph& operator=(Graph&& o) noexcept {
        std::swap(g, o.g);
        return *this;
    }
    ~Graph() { ggpu_graph_free(g); }
    uint64_t n() const { return ggpu_graph_n(g); }
    uint64_t m() const { return ggpu_graph_m(g); }
    bool directed() const { return ggpu_graph_flags(g) & GGPU_FLAG_DIRECTED; }
    uint32_t deg(uint64_t v) const { return ggpu_graph_row_ptr(g)[v + 1] - ggpu_graph_row_ptr(g)[v]; }
};

Graph symmetrized(const Graph& a) {
    ggpu_graph* o = nullptr;
    host_check(ggpu_graph_symmetrize(a.g, &o), "symmetrize");
    return Graph(o);
}

Graph permuted(const Graph& a, const std::vector<uint32_t>& perm) {
    ggpu_graph* o = nullptr;
    host_check(ggpu_graph_permute(a.g, perm.data(), &o), "permute");
    return Graph(o);
}

Graph transposed(const Graph& a) {
    ggpu_graph* o = nullptr;
    host_check(ggpu_graph_transpose(a.g, &o), "transpose");
    return Graph(o);
}

struct Device {
    ggpu_ctx* ctx = nullptr;
    std::string name;
    Device() {
        if (ggpu_ctx_create(0, &ctx) != GGPU_OK) throw Failure("no usable gfx950 device (ggpu_ctx_create failed)");
        char buf[256];
        int cus = 0;
        uint64_t hbm = 0;
        ggpu_device_info(ctx, buf, sizeof buf, &cus, &hbm);
        name = buf;
    }
    ~Device() { ggpu_ctx_destroy(ctx); }
    void check(int rc, const char* what) const {
        if (rc != GGPU_OK) throw Failure(std::string(what) + ": " + ggpu_last_error(ctx));
    }
};

Graph load_graph(const std::string& spec, std::string& dataset) {
    ggpu_graph* g = nullptr;
    if (spec.rfind("synth:", 0) == 0) {
        dataset = spec.substr(6);
        if (ggpu_gen_named(dataset.c_str(), &g) != GGPU_OK) throw Failure("unknown synthetic graph '" + dataset + "'");
        return Graph(g);
    }
    const size_t s = spec.find_last_of('/');
    dataset = spec.substr(s == std::string::npos ? 0 : s + 1);
    dataset = dataset.substr(0, dataset.find('.'));
    host_check(ggpu_graph_load(spec.c_str(), &g), "load");
    return Graph(g);
}

std::ofstream open_csv(const std::string& path) {
    std::ofstream csv;
    if (path.empty()) return csv;
    const bool fresh = !file_exists(path);
    csv.open(path, std::ios::app);
    if (fresh)
        csv << "timestamp,git_commit,chip,dataset,n,m,ordering,algo,variant,params,source,trial,"
               "time_ms_e2e,time_ms_gpu,gteps,iters,validated,load1\n";
    return csv;
}

struct RunInfo {
    std::string ts = timestamp(), commit = run_cmd("git rev-parse --short HEAD 2>/dev/null"), chip;
    double load = load1();
};

// --ordering original|random|degree|rcm (main.cpp:179-201): perm[old] = new
// computed on the symmetrized graph (empty for "original"); --order-seed for
// "random"; --locality-csv keeps one row per dataset/ordering with the
// relabeled graph's locality.
std::vector<uint32_t> ordering_perm(const Args& a, const Graph& sym, const std::string& dataset) {
    const std::string ordering = a.get("ordering", "original");
    std::vector<uint32_t> perm;
    const double t0 = now_ms();
    if (ordering != "original") {
        perm.resize(sym.n());
        host_check(ggpu_order_perm(sym.g, ordering.c_str(), (uint64_t)a.num("order-seed", 1), perm.data()),
                   "ordering");
    }
    const double order_ms = now_ms() - t0;
    const std::string loc_csv = a.get("locality-csv", "");
    if (!loc_csv.empty()) {
        double gap = 0, same = 0;
        if (perm.empty()) {
            host_check(ggpu_graph_locality(sym.g, &gap, &same), "locality");
        } else {
            Graph q = permuted(sym, perm);
            host_check(ggpu_graph_locality(q.g, &gap, &same), "locality");
        }
        char row[512];
        std::snprintf(row, sizeof row, "%s/%s,%s,%s,%llu,%llu,%.2f,%.6f,%.0f", dataset.c_str(), ordering.c_str(),
                      dataset.c_str(), ordering.c_str(), (unsigned long long)sym.n(), (unsigned long long)sym.m(), gap,
                      same, order_ms);
        cli::csv_upsert(loc_csv, "key,dataset,ordering,n,m_sym,mean_gap,same_block64,order_ms",
                        dataset + "/" + ordering, {row});
        std::printf("locality (%s): mean |u-v| %.1f, same 64-block %.2f%%, ordering took %.0f ms\n", ordering.c_str(),
                    gap, 100 * same, order_ms);
    }
    return perm;
}

uint64_t comp_edges(const Graph& g, const std::vector<uint32_t>& dist) {
    uint64_t s = 0;
    for (uint64_t v = 0; v < g.n(); ++v)
        if (dist[v] != GGPU_INF) s += g.deg(v);
    return s / 2;
}

int cpu_schedule(const std::string& s) {
    if (s == "static") return 0;
    if (s == "dynamic,64" || s == "dynamic") return 1;
    if (s == "guided") return 2;
    throw Failure("unknown schedule '" + s + "' (static | dynamic,64 | guided)");
}

std::vector<float> read_policy(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    std::vector<float> p(GGPU_MLP_PARAMS);
    if (!f.read(reinterpret_cast<char*>(p.data()), p.size() * sizeof(float)))
        throw Failure("cannot read " + std::to_string(GGPU_MLP_PARAMS) + " floats from " + path);
    return p;
}

void set_policy(const Args& a, const Device& d, ggpu_bfs_opts& o, int slot) {
    const std::string pol = a.get("policy", "heuristic");
    if (pol == "heuristic") {
        o.policy = GGPU_POLICY_HEURISTIC;
    } else if (pol == "learned") {
        o.policy = GGPU_POLICY_LEARNED;
        if (a.has("policy-file")) {
            std::vector<float> p = read_policy(a.get("policy-file", ""));
            d.check(ggpu_policy_set(d.ctx, slot, p.data(), (uint32_t)p.size()), "policy");
        }
    } else {
        throw Failure("unknown --policy '" + pol + "' (heuristic | learned)");
    }
}

// ------------------------------------------------------------------- BFS
const char* kBfsVariants[] = {"cpu_serial", "cpu_omp", "cpu_do",   "gpu_vertex", "gpu_naive",
                              "gpu_binned", "gpu_edge", "gpu_do",   "hybrid"};

int run_bfs(const Args& a, const Graph& stored, const std::string& dataset) {
    const std::string variant = a.get("variant", "cpu_serial");
    if (std::find(std::begin(kBfsVariants), std::end(kBfsVariants), variant) == std::end(kBfsVariants)) {
        std::fprintf(stderr, "unknown BFS variant '%s'. Available:", variant.c_str());
        for (auto v : kBfsVariants) std::fprintf(stderr, " %s", v);
        std::fprintf(stderr, "\n");
        return 2;
    }
    const bool gpu = variant != "cpu_serial" && variant != "cpu_omp" && variant != "cpu_do";
    const std::string ordering = a.get("ordering", "original");
    ggpu_cpu_bfs_opts copt;
    ggpu_cpu_bfs_opts_default(&copt);
    copt.threads = (int)a.num("threads", 0);
    copt.schedule = cpu_schedule(a.get("schedule", "dynamic,64"));
    copt.alpha = a.real("alpha", 15.0);
    copt.beta = a.real("beta", 18.0);

    const double t_prep = now_ms();
    // ref_g keeps the original labels (sources, the reference BFS, TEPS); g is
    // what the variant runs on, relabeled when --ordering is not original
    Graph g = symmetrized(stored);
    const std::vector<uint32_t> perm = ordering_perm(a, g, dataset);
    const bool relabeled = !perm.empty();
    Graph ref_store;
    if (relabeled) {
        ref_store = std::move(g);
        g = permuted(ref_store, perm);
    }
    const Graph& ref_g = relabeled ? ref_store : g;
    auto to_run = [&](uint32_t s) { return relabeled ? perm[s] : s; };
    std::unique_ptr<Device> dev;
    ggpu_bfs* gb = nullptr;
    ggpu_bfs_opts gopt;
    int gvariant = 0;
    const std::string t_arg = a.get("T", "100000");
    if (gpu) {
        dev = std::make_unique<Device>();
        ggpu_bfs_opts_default(&gopt);
        gopt.tg = (uint32_t)a.num("tg", 256);
        gopt.small_max = (uint32_t)a.num("small-max", 16);
        gopt.large_min = (uint32_t)a.num("large-min", 256);
        gopt.ept = (uint32_t)a.num("ept", 2);
        gopt.alpha = copt.alpha;
        gopt.beta = copt.beta;
        gopt.cpu_threads = (variant == "hybrid" && !a.has("threads")) ? 4 : copt.threads;
        gopt.hybrid_t = t_arg == "inf" ? UINT64_MAX : std::stoull(t_arg);
        const std::string loop = a.get("loop", "host");
        gopt.loop = loop == "host" ? GGPU_LOOP_HOST : loop == "graph" ? GGPU_LOOP_GRAPH
                    : loop == "persistent" ? GGPU_LOOP_PERSISTENT : 99u;
        if (gopt.loop == 99u) throw Failure("unknown --loop '" + loop + "' (host | graph | persistent)");
        set_policy(a, *dev, gopt, GGPU_POLICY_SINGLE);
        dev->check(ggpu_bfs_parse_variant(variant.c_str(), &gvariant), "variant");
        dev->check(ggpu_bfs_create(dev->ctx, g.n(), g.m(), ggpu_graph_row_ptr(g.g), ggpu_graph_col_idx(g.g), &gb),
                   "upload");
    }
    std::unique_ptr<ggpu_bfs, void (*)(ggpu_bfs*)> gb_owner(gb, ggpu_bfs_destroy);
    const double prep_ms = now_ms() - t_prep;

    std::vector<uint32_t> sources;
    if (a.has("source")) {
        sources = {(uint32_t)a.num("source", 0)};
    } else {
        sources.resize((size_t)std::max<long long>(1, a.num("sources", 64)));
        uint32_t k = 0;
        host_check(ggpu_pick_sources(ref_g.g, (uint32_t)sources.size(), (uint64_t)a.num("seed", 1),
                                     a.has("any-component"), sources.data(), &k), "sources");
        sources.resize(k);
    }
    if (sources.empty()) {
        std::fprintf(stderr, "no vertex with degree > 0\n");
        return 1;
    }
    for (uint32_t s : sources)
        if (s >= g.n()) {
            std::fprintf(stderr, "source %u out of range\n", s);
            return 2;
        }

    std::string params;
    if (variant == "cpu_omp" || variant == "cpu_do")
        params = "threads=" + std::to_string(copt.threads > 0 ? copt.threads : omp_get_max_threads());
    if (variant == "cpu_omp")
        params += std::string(" schedule=") + (copt.schedule == 0 ? "static" : copt.schedule == 1 ? "dynamic/64" : "guided");
    char b[160];
    if (variant == "cpu_do") {
        std::snprintf(b, sizeof b, " alpha=%g beta=%g", copt.alpha, copt.beta);
        params += b;
    }
    if (gpu) params = "tg=" + std::to_string(gopt.tg);
    if (variant == "gpu_binned")
        params += " small_max=" + std::to_string(gopt.small_max) + " large_min=" + std::to_string(gopt.large_min);
    if (variant == "gpu_edge" || variant == "gpu_do" || variant == "hybrid") params += " ept=" + std::to_string(gopt.ept);
    if (variant == "gpu_do" || variant == "hybrid") {
        std::snprintf(b, sizeof b, " alpha=%g beta=%g", gopt.alpha, gopt.beta);
        params += b;
    }
    if (variant == "hybrid")
        params += " T=" + t_arg + " threads=" + std::to_string(gopt.cpu_threads > 0 ? gopt.cpu_threads : omp_get_max_threads());
    if (gpu && gopt.loop != GGPU_LOOP_HOST) params += gopt.loop == GGPU_LOOP_GRAPH ? " loop=graph" : " loop=persistent";
    if (gpu && gopt.policy == GGPU_POLICY_LEARNED) params += " policy=learned";

    const int warmup = (int)a.num("warmup", 2), trials = (int)a.num("trials", 5);
    const int source_warmup = (int)a.num("source-warmup", 1);
    const bool validate = a.has("validate");
    std::printf("%s: n %llu, symmetrized m %llu (%llu undirected edges), prep %.0f ms\n", dataset.c_str(),
                (unsigned long long)g.n(), (unsigned long long)g.m(), (unsigned long long)(g.m() / 2), prep_ms);
    std::printf("bfs %s [%s], %zu source(s), %d warm-up + %d trial(s)%s\n", variant.c_str(), params.c_str(),
                sources.size(), warmup, trials, validate ? ", validating" : "");

    std::vector<uint32_t> dist(g.n()), ref(g.n());
    std::vector<ggpu_level> trace;
    // One run; GPU results stay on the device until fetched.
    auto run = [&](uint32_t s, bool fetch, std::vector<ggpu_level>* tr) {
        size_t len = 0;
        if (tr) tr->assign(g.n() + 2, ggpu_level{});
        if (gpu) {
            dev->check(ggpu_bfs_run(gb, gvariant, s, &gopt, tr ? tr->data() : nullptr, tr ? tr->size() : 0, &len),
                       "bfs");
            if (fetch) dev->check(ggpu_bfs_fetch_dist(gb, dist.data()), "fetch");
        } else {
            host_check(ggpu_cpu_bfs(g.g, variant.c_str(), s, &copt, dist.data(), tr ? tr->data() : nullptr,
                                    tr ? tr->size() : 0, &len), "bfs");
        }
        if (tr) tr->resize(std::min(len, tr->size()));
    };
    for (int w = 0; w < warmup; ++w) run(to_run(sources[w % sources.size()]), false, nullptr);

    std::ofstream csv = open_csv(a.get("csv", ""));
    RunInfo info;
    info.chip = no_commas(gpu ? dev->name : cpu_name());
    std::vector<double> all, all_gpu, per_source;
    for (uint32_t s : sources) {
        host_check(ggpu_cpu_bfs(ref_g.g, "cpu_serial", s, &copt, ref.data(), nullptr, 0, nullptr), "reference");
        const uint64_t ce = comp_edges(ref_g, ref);
        const uint32_t rs = to_run(s);
        for (int w = 0; w < source_warmup && trials > 0; ++w) run(rs, false, nullptr);
        std::vector<double> times;
        for (int t = 0; t < trials; ++t) {
            const double t0 = now_ms();
            run(rs, false, nullptr);
            const double ms = now_ms() - t0;
            const double gms = gpu ? ggpu_bfs_last_gpu_ms(gb) : -1;
            if (validate) {
                if (gpu) dev->check(ggpu_bfs_fetch_dist(gb, dist.data()), "fetch");
                // dist (run ids) equals ref (original ids) after v -> perm[v]
                for (uint64_t v = 0; v < g.n(); ++v) {
                    const uint32_t got = dist[relabeled ? perm[v] : v];
                    if (got != ref[v]) {
                        std::fprintf(stderr, "VALIDATION FAILED: %s [%s] source %u trial %d: vertex %llu: got %u, expected %u\n",
                                     variant.c_str(), ordering.c_str(), s, t, (unsigned long long)v, got, ref[v]);
                        return 1;
                    }
                }
            }
            times.push_back(ms);
            all.push_back(ms);
            if (gms >= 0) all_gpu.push_back(gms);
            if (csv.is_open()) {
                char gf[32] = "";
                if (gms >= 0) std::snprintf(gf, sizeof gf, "%.4f", gms);
                char row[1024];
                std::snprintf(row, sizeof row, "%s,%s,%s,%s,%llu,%llu,%s,bfs,%s,%s,%u,%d,%.4f,%s,%.6f,,%d,%.2f\n",
                              info.ts.c_str(), info.commit.c_str(), info.chip.c_str(), dataset.c_str(),
                              (unsigned long long)g.n(), (unsigned long long)g.m(), ordering.c_str(), variant.c_str(),
                              params.c_str(), s,
                              t, ms, gf, ce / (ms * 1e6), validate ? 1 : 0, info.load);
                csv << row;
            }
        }
        if (trials > 0) per_source.push_back(ce / (median(times) * 1e6));
    }
    double hsum = 0;
    size_t hcount = 0;
    for (double x : per_source)
        if (x > 0) {
            hsum += 1.0 / x;
            ++hcount;
        }
    if (!all.empty()) {
        char gp[64] = "";
        if (!all_gpu.empty()) std::snprintf(gp, sizeof gp, " | GPU-only median %.3f", median(all_gpu));
        std::printf("time ms: median %.3f, min %.3f%s | GTEPS harmonic mean %.4f%s\n", median(all),
                    *std::min_element(all.begin(), all.end()), gp, hcount ? hcount / hsum : 0.0,
                    validate ? " | validated OK" : "");
    }
    const std::string trace_path = a.get("trace", "");
    if (!trace_path.empty()) {
        const bool fresh = !file_exists(trace_path);
        std::ofstream tr(trace_path, std::ios::app);
        if (fresh) tr << "dataset,ordering,variant,params,source,level,frontier,edges,direction,device,time_ms\n";
        for (uint32_t s : sources) {
            run(to_run(s), false, &trace);
            for (auto& L : trace) {
                char row[512];
                std::snprintf(row, sizeof row, "%s,%s,%s,%s,%u,%u,%llu,%llu,%c,%c,%.4f\n", dataset.c_str(),
                              ordering.c_str(), variant.c_str(), params.c_str(), s, L.level, (unsigned long long)L.frontier,
                              (unsigned long long)L.edges, L.direction, L.device, L.time_ms);
                tr << row;
            }
        }
        std::printf("trace written to %s\n", trace_path.c_str());
    }
    const std::string dump = a.get("dump-dist", "");
    if (!dump.empty()) {
        run(to_run(sources.back()), true, nullptr);
        std::ofstream out(dump, std::ios::binary);
        out.write(reinterpret_cast<const char*>(dist.data()), g.n() * sizeof(uint32_t));
    }
    return 0;
}

// ------------------------------------------------------- batched BFS (new)
int run_msbfs(const Args& a, const Graph& stored, const std::string& dataset) {
    const std::string border = a.get("batch-order", "auto");
    uint32_t batch_order = GGPU_BATCH_ORDER_AUTO;
    if (border == "degree") batch_order = GGPU_BATCH_ORDER_DEGREE;
    else if (border == "given") batch_order = GGPU_BATCH_ORDER_GIVEN;
    else if (border != "auto") {
        std::fprintf(stderr, "unknown --batch-order '%s'. Available: auto degree given\n", border.c_str());
        return 2;
    }
    Graph g = symmetrized(stored);
    const std::string ordering = a.get("ordering", "original");
    const std::vector<uint32_t> perm = ordering_perm(a, g, dataset);
    const bool relabeled = !perm.empty();
    Graph ref_store;
    if (relabeled) {
        ref_store = std::move(g);
        g = permuted(ref_store, perm);
    }
    const Graph& ref_g = relabeled ? ref_store : g;  // original labels
    Device dev;
    ggpu_bfs_opts o;
    ggpu_bfs_opts_default(&o);
    o.alpha = a.real("alpha", 15.0);
    o.beta = a.real("beta", 18.0);
    set_policy(a, dev, o, GGPU_POLICY_BATCHED);
    o.batch_order = batch_order;
    ggpu_bfs* gb = nullptr;
    dev.check(ggpu_bfs_create(dev.ctx, g.n(), g.m(), ggpu_graph_row_ptr(g.g), ggpu_graph_col_idx(g.g), &gb), "upload");
    std::unique_ptr<ggpu_bfs, void (*)(ggpu_bfs*)> owner(gb, ggpu_bfs_destroy);
    std::vector<uint32_t> sources((size_t)std::max<long long>(1, a.num("sources", 64)));
    uint32_t k = 0;
    host_check(ggpu_pick_sources(ref_g.g, (uint32_t)sources.size(), (uint64_t)a.num("seed", 1),
                                 a.has("any-component"), sources.data(), &k), "sources");
    sources.resize(k);
    std::vector<uint32_t> run_sources(sources);  // run ids
    if (relabeled)
        for (auto& x : run_sources) x = perm[x];
    if (sources.empty()) {
        std::fprintf(stderr, "no vertex with degree > 0\n");
        return 1;
    }
    const uint32_t batch = (uint32_t)std::min<long long>(256, std::max<long long>(1, a.num("batch", 64)));
    const int warmup = (int)a.num("warmup", 2), trials = (int)a.num("trials", 5);
    const bool validate = a.has("validate");
    const std::string params =
        "batch=" + std::to_string(batch) + " policy=" + a.get("policy", "heuristic") + " batch_order=" + border;
    std::printf("%s: n %llu, symmetrized m %llu\nmsbfs [%s], %zu sources, %d warm-up + %d trial(s)%s\n",
                dataset.c_str(), (unsigned long long)g.n(), (unsigned long long)g.m(), params.c_str(), sources.size(),
                warmup, trials, validate ? ", validating" : "");
    ggpu_cpu_bfs_opts copt;
    ggpu_cpu_bfs_opts_default(&copt);
    std::vector<uint32_t> ref(g.n());
    std::vector<uint64_t> ce(sources.size());
    for (size_t i = 0; i < sources.size(); ++i) {
        host_check(ggpu_cpu_bfs(ref_g.g, "cpu_serial", sources[i], &copt, ref.data(), nullptr, 0, nullptr),
                   "reference");
        ce[i] = comp_edges(ref_g, ref);
    }
    std::ofstream csv = open_csv(a.get("csv", ""));
    RunInfo info;
    info.chip = no_commas(dev.name);
    std::vector<double> all, all_gpu;
    std::vector<uint32_t> dist;
    double total_edges = 0;
    for (uint64_t x : ce) total_edges += (double)x;
    for (int t = -warmup; t < trials; ++t) {
        double ms = 0, gms = 0;
        for (size_t b0 = 0; b0 < sources.size(); b0 += batch) {
            const uint32_t kb = (uint32_t)std::min<size_t>(batch, sources.size() - b0);
            ggpu_msbfs_result r{};
            const double t0 = now_ms();
            dev.check(ggpu_msbfs_run(gb, run_sources.data() + b0, kb, &o, &r, nullptr, 0, nullptr), "msbfs");
            ms += now_ms() - t0;
            gms += r.gpu_ms;
            if (validate && t >= 0) {
                dist.resize((size_t)kb * g.n());
                dev.check(ggpu_msbfs_fetch_dist(gb, dist.data()), "fetch");
                for (uint32_t i = 0; i < kb; ++i) {
                    host_check(ggpu_cpu_bfs(ref_g.g, "cpu_serial", sources[b0 + i], &copt, ref.data(), nullptr, 0,
                                            nullptr),
                               "reference");
                    const uint32_t* d = dist.data() + (size_t)i * g.n();
                    for (uint64_t v = 0; v < g.n(); ++v)
                        if (d[relabeled ? perm[v] : v] != ref[v]) {
                            std::fprintf(stderr, "VALIDATION FAILED: msbfs [%s] source %u (batch position %u) trial %d\n",
                                         ordering.c_str(), sources[b0 + i], i, t);
                            return 1;
                        }
                }
            }
        }
        if (t < 0) continue;
        all.push_back(ms);
        all_gpu.push_back(gms);
        if (csv.is_open()) {
            char row[1024];
            std::snprintf(row, sizeof row, "%s,%s,%s,%s,%llu,%llu,%s,msbfs,gpu_msbfs,%s,%u,%d,%.4f,%.4f,%.6f,,%d,%.2f\n",
                          info.ts.c_str(), info.commit.c_str(), info.chip.c_str(), dataset.c_str(),
                          (unsigned long long)g.n(), (unsigned long long)g.m(), ordering.c_str(), params.c_str(),
                          sources[0], t, ms, gms,
                          total_edges / (ms * 1e6), validate ? 1 : 0, info.load);
            csv << row;
        }
    }
    if (!all.empty())
        std::printf("time ms (all %zu sources): median %.3f, min %.3f | GPU-only median %.3f | aggregate GTEPS %.3f%s\n",
                    sources.size(), median(all), *std::min_element(all.begin(), all.end()), median(all_gpu),
                    total_edges / (median(all) * 1e6), validate ? " | validated OK" : "");
    return 0;
}

// -------------------------------------------------------------- PageRank
const char* kPrVariants[] = {"cpu_serial", "cpu_omp", "gpu_push", "gpu_pull_scalar", "gpu_pull_vector", "gpu_pull_adaptive",
                             "gpu_pull_gs"};

// cpu::gsNoWorseThanJacobi (src/cpu/pagerank_cpu.hpp:80-84)
bool gs_no_worse(double l1, double mx, double sum, double jl1, double jmx) {
    const double budget = std::max(1e-4, jl1);
    return l1 <= budget && std::fabs(sum - 1.0) <= budget && mx <= std::max(1e-3, jmx * (1 + 1e-4));
}

struct Check {
    double l1 = 0, mx = 0, sum = 0;
    bool ok() const { return l1 < 1e-4 && mx < 1e-3 && std::fabs(sum - 1.0) < 1e-4; }
};

Check check(const std::vector<double>& ref, const std::vector<float>& got) {
    Check c;
    ggpu_check_ranks(ref.size(), ref.data(), got.data(), &c.l1, &c.mx, &c.sum);
    return c;
}

int run_pr(const Args& a, const Graph& stored, const std::string& dataset) {
    const std::string variant = a.get("variant", "cpu_serial");
    if (std::find(std::begin(kPrVariants), std::end(kPrVariants), variant) == std::end(kPrVariants)) {
        std::fprintf(stderr, "unknown PageRank variant '%s'. Available:", variant.c_str());
        for (auto v : kPrVariants) std::fprintf(stderr, " %s", v);
        std::fprintf(stderr, "\n");
        return 2;
    }
    const bool gpu = variant.rfind("gpu_", 0) == 0;
    const std::string ordering = a.get("ordering", "original");
    const double tol = a.real("tol", 0);
    const int iters = (int)a.num("iters", tol > 0 ? 1000 : 20);
    const int threads = (int)a.num("threads", 0);
    const double t_prep = now_ms();
    const uint64_t n = stored.n(), m = stored.m();
    std::vector<uint32_t> outdeg(n);
    for (uint64_t v = 0; v < n; ++v) outdeg[v] = stored.deg(v);
    Graph in_owned;
    const Graph* in = &stored;
    if (stored.directed()) {
        in_owned = transposed(stored);
        in = &in_owned;
    }
    const bool validate = a.has("validate");
    const bool in_place = a.has("inplace");
    const bool gs_check = variant == "gpu_pull_gs" || in_place;
    const double ref_tol = a.real("ref-tol", tol);
    std::vector<double> ref, ref_conv;
    int ref_iters = 0;
    Check jacobi;
    double dummy = 0;
    if (validate) {
        ref.resize(n);
        const double rt = gs_check ? ref_tol : tol;
        const int ri = gs_check && ref_tol != tol ? (ref_tol > 0 ? 1000 : iters) : iters;
        host_check(ggpu_cpu_pagerank(in->g, outdeg.data(), 0.85, ri, rt, 0, 0, ref.data(), &ref_iters, &dummy), "ref");
        if (gs_check) {
            ref_conv.resize(n);
            int ci = 0;
            host_check(ggpu_cpu_pagerank(in->g, outdeg.data(), 0.85, 1000, 1e-10, 0, 0, ref_conv.data(), &ci, &dummy),
                       "ref");
            // the Jacobi reference's own distance from convergence
            const std::vector<float> rf(ref.begin(), ref.end());
            ggpu_check_ranks(n, ref_conv.data(), rf.data(), &jacobi.l1, &jacobi.mx, &jacobi.sum);
        }
    }
    // --ordering: perm from the symmetrized graph (as for BFS) relabels the
    // directed graph the variant runs on; the reference above keeps the
    // original labels and ranks are mapped back (rank of v = run[perm[v]])
    std::vector<uint32_t> perm;
    {
        Graph sym = symmetrized(stored);
        perm = ordering_perm(a, sym, dataset);
    }
    const bool relabeled = !perm.empty();
    Graph run_graph, run_in_owned;
    const Graph* run_in = in;
    std::vector<uint32_t> run_outdeg = outdeg;
    if (relabeled) {
        run_graph = permuted(stored, perm);
        if (run_graph.directed()) {
            run_in_owned = transposed(run_graph);
            run_in = &run_in_owned;
        } else {
            run_in = &run_graph;
        }
        for (uint64_t v = 0; v < n; ++v) run_outdeg[perm[v]] = outdeg[v];
    }
    std::unique_ptr<Device> dev;
    ggpu_pr* pr = nullptr;
    ggpu_pr_opts po;
    int pvariant = 0;
    if (gpu) {
        dev = std::make_unique<Device>();
        ggpu_pr_opts_default(&po);
        po.iters = iters;
        po.tol = tol;
        po.tol_rel = a.real("tol-rel", 0);
        po.tg = (uint32_t)a.num("tg", 256);
        po.small_max = (uint32_t)a.num("small-max", 16);
        po.large_min = (uint32_t)a.num("large-min", 256);
        po.in_place = in_place;
        dev->check(ggpu_pr_parse_variant(variant.c_str(), &pvariant), "variant");
        dev->check(ggpu_pr_create(dev->ctx, n, run_in->m(), ggpu_graph_row_ptr(run_in->g),
                                  ggpu_graph_col_idx(run_in->g), run_outdeg.data(), &pr), "upload");
        if (variant == "gpu_pull_adaptive") dev->check(ggpu_pr_prepare_adaptive(pr, po.small_max, po.large_min), "prepare");
        if (variant == "gpu_pull_gs") dev->check(ggpu_pr_prepare_gs(pr, (uint32_t)a.num("gs-pairs", 2)), "prepare");
    }
    std::unique_ptr<ggpu_pr, void (*)(ggpu_pr*)> pr_owner(pr, ggpu_pr_destroy);
    const double prep_ms = now_ms() - t_prep;

    char b[64];
    std::snprintf(b, sizeof b, "%g", tol);
    std::string params = tol > 0 ? std::string("tol=") + b + " cap=" + std::to_string(iters) : "iters=" + std::to_string(iters);
    if (variant == "cpu_omp") params += " threads=" + std::to_string(threads > 0 ? threads : omp_get_max_threads());
    if (gpu) params += " tg=" + std::to_string(po.tg);
    if (gpu && po.tol_rel > 0) {
        std::snprintf(b, sizeof b, " tol_rel=%g", po.tol_rel);
        params += b;
    }
    if (in_place) params += " inplace";
    if (variant == "gpu_pull_gs") {
        int colored = 0;
        uint64_t sets = 0, deferred = 0;
        dev->check(ggpu_pr_gs_info(pr, &colored, &sets, &deferred), "gs info");
        params += colored ? " order=colored sets=" + std::to_string(sets)
                          : " order=natural gs_pairs=" + std::to_string(a.num("gs-pairs", 2)) +
                                " deferred=" + std::to_string(deferred);
    }
    if (variant == "gpu_pull_adaptive")
        params += " small_max=" + std::to_string(po.small_max) + " large_min=" + std::to_string(po.large_min);
    std::printf("%s: n %llu, m %llu (directed=%d), ordering %s, prep %.0f ms\npr %s [%s]\n", dataset.c_str(),
                (unsigned long long)n, (unsigned long long)m, (int)stored.directed(), ordering.c_str(), prep_ms,
                variant.c_str(), params.c_str());

    std::vector<double> rcpu(n);
    std::vector<float> ranks(n);
    double last_delta = -1, last_rel = -1;
    int floor_stop = 0;
    auto run_once = [&]() -> std::pair<int, double> {
        if (!gpu) {
            int it = 0;
            host_check(ggpu_cpu_pagerank(run_in->g, run_outdeg.data(), 0.85, iters, tol, variant == "cpu_omp", threads,
                                         rcpu.data(), &it, &last_delta), "pagerank");
            return {it, -1.0};
        }
        ggpu_pr_result r{};
        dev->check(ggpu_pr_run(pr, pvariant, &po, &r), "pagerank");
        last_delta = r.last_delta;
        last_rel = r.last_max_rel;
        floor_stop = r.floor_stop;
        return {r.iters, r.gpu_ms};
    };
    std::vector<float> run_ranks(n);
    auto current = [&]() {  // ranks in the original labels
        if (gpu) dev->check(ggpu_pr_fetch_ranks(pr, run_ranks.data()), "fetch");
        else
            for (uint64_t v = 0; v < n; ++v) run_ranks[v] = (float)rcpu[v];
        for (uint64_t v = 0; v < n; ++v) ranks[v] = run_ranks[relabeled ? perm[v] : v];
        return ranks;
    };
    const int warmup = (int)a.num("warmup", 2), trials = (int)a.num("trials", 5);
    for (int w = 0; w < warmup; ++w) run_once();
    std::ofstream csv = open_csv(a.get("csv", ""));
    RunInfo info;
    info.chip = no_commas(gpu ? dev->name : cpu_name());
    std::vector<double> times, gtimes;
    int it_last = 0;
    Check worst;
    for (int t = 0; t < trials; ++t) {
        const double t0 = now_ms();
        auto [it, gms] = run_once();
        const double ms = now_ms() - t0;
        it_last = it;
        if (validate) {
            const Check c = check(gs_check ? ref_conv : ref, current());
            const bool ok = !gs_check ? c.ok() : tol > 0 ? gs_no_worse(c.l1, c.mx, c.sum, jacobi.l1, jacobi.mx)
                                                         : std::fabs(c.sum - 1.0) < 1e-4;
            if (!ok) {
                std::fprintf(stderr, "VALIDATION FAILED: %s trial %d: L1 %.6g, max rel (top-1000) %.6g, sum %.8f\n",
                             variant.c_str(), t, c.l1, c.mx, c.sum);
                return 1;
            }
            if (c.l1 >= worst.l1) worst = c;
        }
        times.push_back(ms);
        if (gms >= 0) gtimes.push_back(gms);
        if (csv.is_open()) {
            char gf[32] = "";
            if (gms >= 0) std::snprintf(gf, sizeof gf, "%.4f", gms);
            char row[1024];
            std::snprintf(row, sizeof row, "%s,%s,%s,%s,%llu,%llu,%s,pr,%s,%s,,%d,%.4f,%s,%.6f,%d,%d,%.2f\n",
                          info.ts.c_str(), info.commit.c_str(), info.chip.c_str(), dataset.c_str(),
                          (unsigned long long)n, (unsigned long long)m, ordering.c_str(), variant.c_str(),
                          params.c_str(), t, ms, gf,
                          (double)m * it / (ms * 1e6), it, validate ? 1 : 0, info.load);
            csv << row;
        }
    }
    if (!times.empty()) {
        const double med = median(times), per = med / std::max(1, it_last);
        std::printf("time ms: median %.3f, min %.3f | %d iterations, %.4f ms/iteration", med,
                    *std::min_element(times.begin(), times.end()), it_last, per);
        if (!gtimes.empty()) std::printf(" (GPU-only %.4f)", median(gtimes) / std::max(1, it_last));
        if (tol > 0)
            std::printf(" | final L1 change %.3g%s", last_delta,
                        last_delta < tol ? "" : floor_stop ? " (stopped at float precision floor)" : " (NOT converged)");
        if (tol > 0 && last_rel >= 0) std::printf(", max rel change %.3g", last_rel);
        if (variant != "cpu_serial") {
            const double bytes = 4.0 * (n + 1) + 8.0 * m + 4.0 * n;
            std::printf(" | eff. %.1f GB/s, %.3f Ggathers/s", bytes / (per * 1e6), m / (per * 1e6));
        }
        if (validate && gs_check && tol <= 0) {
            ggpu_pr_opts conv = po;
            conv.tol = 0;
            conv.iters = 200;
            ggpu_pr_result r{};
            dev->check(ggpu_pr_run(pr, pvariant, &conv, &r), "pagerank");
            const Check c = check(ref_conv, current());
            if (!c.ok()) {
                std::printf("\n");
                std::fprintf(stderr, "VALIDATION FAILED: %s after 200 sweeps vs converged reference: L1 %.6g, max rel "
                             "(top-1000) %.6g, sum %.8f\n", variant.c_str(), c.l1, c.mx, c.sum);
                return 1;
            }
            std::printf(" | validated OK (200-sweep run vs converged ref: L1 %.2e, top-1000 rel %.2e)", c.l1, c.mx);
        } else if (validate) {
            std::printf(" | validated OK (ref %d iters; worst L1 %.2e, top-1000 rel %.2e, sum-1 %.1e)", ref_iters,
                        worst.l1, worst.mx, worst.sum - 1.0);
        }
        std::printf("\n");
    }
    const std::string dump = a.get("dump-ranks", "");
    if (!dump.empty()) {
        if (trials == 0) run_once();
        const std::vector<float> r = current();
        std::vector<double> out(r.begin(), r.end());
        std::ofstream o(dump, std::ios::binary);
        o.write(reinterpret_cast<const char*>(out.data()), n * sizeof(double));
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        Args a(argc, argv);
        if (!a.has("graph")) {
            std::fprintf(stderr,
                         "usage: gbench --graph <file.bin | synth:NAME> --algo bfs|pr|msbfs --variant V [options]\n"
                         "       (see the comment at the top of gbench.cpp)\n");
            return 2;
        }
        std::string dataset;
        Graph g = load_graph(a.get("graph", ""), dataset);
        if (a.has("save-graph")) {
            host_check(ggpu_graph_save(g.g, a.get("save-graph", "").c_str()), "save");
            return 0;
        }
        const std::string algo = a.get("algo", "bfs");
        if (algo == "bfs") return run_bfs(a, g, dataset);
        if (algo == "pr") return run_pr(a, g, dataset);
        if (algo == "msbfs") return run_msbfs(a, g, dataset);
        std::fprintf(stderr, "unknown --algo '%s' (bfs | pr | msbfs)\n", algo.c_str());
        return 2;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "gbench: %s\n", e.what());
        return 1;
    }
}
