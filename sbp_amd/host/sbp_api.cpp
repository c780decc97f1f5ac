// C++ drop-in surface (include/sbp/sbp.hpp) over the C-ABI (include/sbp_amd.h): Graph, Blockmodel,
// BlockmodelTriplet, the agglomerative driver and the namespaced entry points.  Only C-ABI calls reach
// the device from here.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <fstream>
#include <mutex>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <unordered_set>

#include "../../include/sbp/sbp.hpp"

namespace sbp_amd {

static std::mutex g_mu;
static sbp_ctx g_ctx = nullptr;
static int g_ctx_device = -1;

void throw_on(int status, sbp_ctx ctx, const char *what) {
    if (status == SBP_OK) return;
    std::string msg = std::string(what) + ": " + (ctx ? sbp_last_error(ctx) : "no context");
    if (status == SBP_E_ARG) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

sbp_ctx context(const Args &a) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ctx || g_ctx_device != a.device) {
        if (g_ctx) sbp_destroy(g_ctx);
        g_ctx = nullptr;
        int st = sbp_create(&g_ctx, a.device, a.seed);
        if (st != SBP_OK) throw std::runtime_error("sbp_create failed: no usable HIP device");
        g_ctx_device = a.device;
    }
    sbp_set_seed(g_ctx, a.seed);
    return g_ctx;
}

Args &args() { return ::args; }

uint32_t &epoch() {
    static uint32_t e = 0;
    return e;
}

sbp_params to_params(const Args &a) {
    sbp_params p;
    if (a.algorithm == "async_gibbs" || a.algorithm == "async_gibbs_load_balanced") p.algorithm = 0;
    else if (a.algorithm == "metropolis_hastings") p.algorithm = 1;
    else if (a.algorithm == "hybrid_mcmc") p.algorithm = 2;
    else throw std::invalid_argument("unknown algorithm " + a.algorithm);
    p.batches = a.batches;
    p.greedy = a.greedy ? 1 : 0;
    p.approximate = a.approximate ? 1 : 0;
    p.ordered = a.ordered ? 1 : 0;
    p.asynciterations = a.asynciterations;
    p.mh_percent = a.mh_percent;
    p.seed = a.seed;
    p.max_iterations = 100;
    p.threads = a.threads;
    p.nonparametric = a.nonparametric ? 1 : 0;
    p.degreecorrected = a.degreecorrected ? 1 : 0;
    if (a.nonparametric && a.undirected)
        throw std::invalid_argument("the nonparametric description length is built for directed graphs");
    return p;
}

static thread_local sbp_stats g_stats;

}  // namespace sbp_amd

using namespace sbp_amd;

// ------------------------------------------------------------------ Graph

Graph::Graph(long num_vertices, std::vector<int64_t> out_off, std::vector<int32_t> out_adj,
             std::vector<int64_t> in_off, std::vector<int32_t> in_adj)
    : N_(num_vertices), out_off_(std::move(out_off)), in_off_(std::move(in_off)), out_adj_(std::move(out_adj)),
      in_adj_(std::move(in_adj)) {
    if ((long)out_off_.size() != N_ + 1 || (long)in_off_.size() != N_ + 1)
        throw std::invalid_argument("Graph: offsets must have N+1 entries");
}

Graph Graph::from_edges(long N, const std::vector<std::pair<long, long>> &edges) {
    std::vector<int64_t> oo(N + 1, 0), io(N + 1, 0);
    for (auto &e : edges) {
        if (e.first < 0 || e.first >= N || e.second < 0 || e.second >= N)
            throw std::invalid_argument("Graph::from_edges: vertex id out of range");
        oo[e.first + 1]++;
        io[e.second + 1]++;
    }
    for (long v = 0; v < N; ++v) {
        oo[v + 1] += oo[v];
        io[v + 1] += io[v];
    }
    std::vector<int32_t> oa(edges.size()), ia(edges.size());
    std::vector<int64_t> po(oo.begin(), oo.end() - 1), pi(io.begin(), io.end() - 1);
    for (auto &e : edges) {  // file order within each list (utils::insert)
        oa[po[e.first]++] = (int32_t)e.second;
        ia[pi[e.second]++] = (int32_t)e.first;
    }
    return Graph(N, std::move(oo), std::move(oa), std::move(io), std::move(ia));
}

static std::vector<std::vector<std::string>> read_rows(const std::string &path) {
    std::vector<std::vector<std::string>> rows;
    std::ifstream f(path);
    if (!f) return rows;
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream ss(line);
        std::vector<std::string> row;
        std::string tok;
        while (ss >> tok) row.push_back(tok);
        if (!row.empty()) rows.push_back(row);
    }
    return rows;
}

Graph Graph::load(const Args &a) {
    std::vector<std::pair<long, long>> edges;
    long N = 0;
    auto rows = read_rows(a.filepath + ".tsv");
    bool mtx = false;
    if (rows.empty()) {
        rows = read_rows(a.filepath + ".mtx");
        mtx = true;
    }
    if (rows.empty()) throw std::invalid_argument("Graph::load: no " + a.filepath + ".tsv or .mtx");
    size_t first = 0;
    bool undirected = a.undirected;
    if (mtx || rows[0][0] == "%%MatrixMarket") {
        // Graph::load_matrix_market (src/graph.cpp:75-129)
        if (rows[0].size() > 4 && rows[0][4] == "symmetric") undirected = true;
        while (first < rows.size() && rows[first][0][0] == '%') ++first;
        N = std::stol(rows[first][0]);
        ++first;
    }
    for (size_t i = first; i < rows.size(); ++i) {
        long from = std::stol(rows[i][0]) - 1, to = std::stol(rows[i][1]) - 1;
        N = std::max(N, std::max(from, to) + 1);
        edges.emplace_back(from, to);
        if (undirected && from != to) edges.emplace_back(to, from);
    }
    if (a.noduplicates) {  // utils::read_csv drops repeated lines when --noduplicates (src/utils.cpp:100-120)
        std::vector<std::pair<long, long>> uniq;
        std::vector<std::pair<long, long>> seen(edges);
        std::sort(seen.begin(), seen.end());
        seen.erase(std::unique(seen.begin(), seen.end()), seen.end());
        std::vector<char> used(seen.size(), 0);
        for (auto &e : edges) {
            size_t k = std::lower_bound(seen.begin(), seen.end(), e) - seen.begin();
            if (!used[k]) {
                used[k] = 1;
                uniq.push_back(e);
            }
        }
        edges.swap(uniq);
    }
    Graph g = from_edges(N, edges);
    auto truth = read_rows(a.filepath + "_truePartition.tsv");
    std::vector<long> assignment(N, 0);
    if (!truth.empty()) {
        std::fill(assignment.begin(), assignment.end(), -1);
        for (auto &r : truth) {
            long v = std::stol(r[0]) - 1, c = std::stol(r[1]) - 1;
            if (v >= 0 && v < N) assignment[v] = c;
        }
    }
    g.assignment(assignment);
    return g;
}

std::vector<long> Graph::out_neighbors(long v) const {
    return std::vector<long>(out_adj_.begin() + out_off_[v], out_adj_.begin() + out_off_[v + 1]);
}
std::vector<long> Graph::in_neighbors(long v) const {
    return std::vector<long>(in_adj_.begin() + in_off_[v], in_adj_.begin() + in_off_[v + 1]);
}

std::vector<long> Graph::degrees() const {
    std::vector<long> d(N_);
    for (long v = 0; v < N_; ++v) {
        long self = 0;
        for (int64_t e = out_off_[v]; e < out_off_[v + 1]; ++e)
            if (out_adj_[e] == v) self = 1;
        d[v] = (long)(out_off_[v + 1] - out_off_[v] + in_off_[v + 1] - in_off_[v]) - self;
    }
    return d;
}

sbp_graph Graph::device(sbp_ctx ctx) const {
    if (!dev_) {
        sbp_graph g = nullptr;
        throw_on(sbp_graph_upload(ctx, (int32_t)N_, (int64_t)out_adj_.size(), out_off_.data(), out_adj_.data(),
                                  in_off_.data(), in_adj_.data(), &g),
                 ctx, "sbp_graph_upload");
        dev_ = std::shared_ptr<sbp_graph_s>(g, [](sbp_graph x) { sbp_graph_free(x); });
    }
    return dev_.get();
}

Graph Graph::adopt(sbp_ctx ctx, sbp_graph g, std::shared_ptr<struct sbp_graph_s> keep_alive) {
    int32_t N = 0;
    int64_t E = 0;
    throw_on(sbp_graph_info(g, &N, &E), ctx, "sbp_graph_info");
    std::vector<int64_t> oo(N + 1), io(N + 1);
    std::vector<int32_t> oa(E), ia(E);
    throw_on(sbp_graph_download(ctx, g, oo.data(), oa.data(), io.data(), ia.data()), ctx, "sbp_graph_download");
    Graph out(N, std::move(oo), std::move(oa), std::move(io), std::move(ia));
    out.dev_ = std::shared_ptr<sbp_graph_s>(g, [](sbp_graph x) { sbp_graph_free(x); });
    out.parent_ = std::move(keep_alive);
    return out;
}

// ------------------------------------------------------------------ Blockmodel

static std::shared_ptr<sbp_bm_s> wrap(sbp_bm bm) {
    return std::shared_ptr<sbp_bm_s>(bm, [](sbp_bm x) { sbp_bm_free(x); });
}

Blockmodel::Blockmodel(long num_blocks, const Graph &graph, double rate) {
    if (num_blocks != graph.num_vertices())
        throw std::invalid_argument("Blockmodel(num_blocks, graph, rate) places every vertex in its own block");
    std::vector<long> a(graph.num_vertices());
    std::iota(a.begin(), a.end(), 0L);
    *this = Blockmodel(num_blocks, graph, rate, a);
}

Blockmodel::Blockmodel(long num_blocks, const Graph &graph, double rate, std::vector<long> &assignment) {
    ctx_ = context(sbp_amd::args());
    N_ = graph.num_vertices();
    std::vector<int32_t> a(assignment.begin(), assignment.end());
    sbp_bm bm = nullptr;
    throw_on(sbp_bm_build(ctx_, graph.device(ctx_), (int32_t)num_blocks, a.data(), &bm), ctx_, "sbp_bm_build");
    const Args &args = sbp_amd::args();
    throw_on(sbp_bm_set_model(ctx_, bm, graph.device(ctx_), args.nonparametric ? (args.degreecorrected ? 2 : 1) : 0),
             ctx_, "sbp_bm_set_model");
    bm_ = wrap(bm);
    graph_ = &graph;
    num_blocks_to_merge_ = (long)((double)num_blocks * rate);  // blockmodel.hpp:70-76
    empty = false;
}

Blockmodel Blockmodel::adopt(sbp_ctx ctx, sbp_bm bm, long num_vertices, double overall_entropy,
                             long num_blocks_to_merge) {
    Blockmodel b;
    b.bm_ = wrap(bm);
    b.ctx_ = ctx;
    b.N_ = num_vertices;
    b.overall_entropy_ = overall_entropy;
    b.num_blocks_to_merge_ = num_blocks_to_merge;
    b.empty = false;
    return b;
}

Blockmodel Blockmodel::copy() const {
    if (empty) return Blockmodel();
    sbp_bm out = nullptr;
    throw_on(sbp_bm_copy(ctx_, bm_.get(), &out), ctx_, "sbp_bm_copy");
    Blockmodel b = adopt(ctx_, out, N_, overall_entropy_, 0);  // copy() resets num_blocks_to_merge (blockmodel.cpp:152)
    b.graph_ = graph_;
    return b;
}

long Blockmodel::num_blocks() const {
    if (empty) return 0;
    int32_t B = 0;
    sbp_bm_info(bm_.get(), &B, nullptr);
    return B;
}

static std::vector<long> widen(const std::vector<int32_t> &v) { return std::vector<long>(v.begin(), v.end()); }

std::vector<long> Blockmodel::block_assignment() const {
    std::vector<int32_t> a(N_);
    throw_on(sbp_bm_download(ctx_, bm_.get(), a.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr),
             ctx_, "sbp_bm_download");
    return widen(a);
}

static std::vector<long> block_array(sbp_ctx ctx, sbp_bm bm, int which) {
    int32_t B = 0;
    sbp_bm_info(bm, &B, nullptr);
    std::vector<int32_t> v(B);
    int32_t *p[4] = {nullptr, nullptr, nullptr, nullptr};
    p[which] = v.data();
    throw_on(sbp_bm_download(ctx, bm, nullptr, p[0], p[1], p[2], p[3], nullptr, nullptr, nullptr), ctx,
             "sbp_bm_download");
    return widen(v);
}

std::vector<long> Blockmodel::block_sizes() const { return block_array(ctx_, bm_.get(), 0); }
std::vector<long> Blockmodel::degrees_out() const { return block_array(ctx_, bm_.get(), 1); }
std::vector<long> Blockmodel::degrees_in() const { return block_array(ctx_, bm_.get(), 2); }
std::vector<long> Blockmodel::degrees() const { return block_array(ctx_, bm_.get(), 3); }

bool Blockmodel::validate(const Graph &graph) const {
    int32_t ok = 0;
    throw_on(sbp_bm_validate(ctx_, bm_.get(), graph.device(ctx_), &ok), ctx_, "sbp_bm_validate");
    return ok != 0;
}

double Blockmodel::log_posterior_probability(const Graph &graph) const {
    double m = 0, lp = 0;
    throw_on(sbp_mdl(ctx_, bm_.get(), graph.device(ctx_), &m, &lp), ctx_, "sbp_mdl");
    return lp;
}

// ------------------------------------------------------------------ BlockmodelTriplet

void BlockmodelTriplet::update(Blockmodel &b) {
    long index;
    if (blockmodels_[1].empty) {
        index = 1;
    } else if (b.getOverall_entropy() <= blockmodels_[1].getOverall_entropy()) {
        long old_index = (blockmodels_[1].num_blocks() > b.num_blocks()) ? 0 : 2;
        blockmodels_[old_index] = blockmodels_[1].copy();
        index = 1;
    } else {
        index = (blockmodels_[1].num_blocks() > b.num_blocks()) ? 2 : 0;
    }
    blockmodels_[index] = b.copy();
}

bool BlockmodelTriplet::is_done() {
    if ((!blockmodels_[0].empty && nb(0) - nb(2) == 2) || (blockmodels_[0].empty && nb(1) - nb(2) == 1))
        optimal_num_blocks_found = true;
    return optimal_num_blocks_found;
}

void BlockmodelTriplet::status() const {
    fprintf(stderr, "Number of blocks: %ld %ld %ld%s\n", nb(0), nb(1), nb(2),
            optimal_num_blocks_found ? " (optimal found)" : "");
}

Blockmodel BlockmodelTriplet::get_next_blockmodel(Blockmodel &old) {
    old.setNum_blocks_to_merge(0);
    update(old);
    if (golden_ratio_not_reached()) {
        Blockmodel b = blockmodels_[1].copy();
        b.setNum_blocks_to_merge((long)((double)b.num_blocks() * 0.5));
        if (b.getNum_blocks_to_merge() == 0) optimal_num_blocks_found = true;
        return b;
    }
    if (is_done()) return blockmodels_[1].copy();
    long index;
    if (blockmodels_[0].empty && nb(1) > nb(2)) index = 1;
    else if (nb(0) - nb(1) >= nb(1) - nb(2)) index = 0;
    else index = 1;
    long next = nb(index + 1);
    next += (long)((double)(nb(index) - nb(index + 1)) * 0.618);
    Blockmodel b = blockmodels_[index].copy();
    b.setNum_blocks_to_merge(b.num_blocks() - next);
    return b;
}

// ------------------------------------------------------------------ driver core

namespace sbp_amd {

static Blockmodel merge_step(sbp_ctx ctx, sbp_graph g, Blockmodel &bm, const sbp_params &p, uint32_t &ep,
                             sbp_stats *st) {
    sbp_bm out = nullptr;
    throw_on(sbp_merge_blocks(ctx, bm.handle(), g, ep++, bm.getNum_blocks_to_merge(), p.approximate, &out, st), ctx,
             "sbp_merge_blocks");
    return Blockmodel::adopt(ctx, out, bm.num_vertices(), bm.getOverall_entropy(), bm.getNum_blocks_to_merge());
}

static void mcmc_step(sbp_ctx ctx, sbp_graph g, Blockmodel &bm, const sbp_params &p, bool grnr, uint32_t &ep,
                      sbp_stats *st) {
    double mdl = 0;
    throw_on(sbp_mcmc(ctx, bm.handle(), g, &p, grnr ? 1 : 0, &ep, &mdl, st), ctx, "sbp_mcmc");
    if (bm.num_blocks() > 1) bm.setOverall_entropy(mdl);
}

// sbp::stochastic_block_partition (src/sbp.cpp:16-72) over a context and a device graph.
Blockmodel drive(sbp_ctx ctx, sbp_graph g, int32_t N, const sbp_params &p, sbp_stats *st) {
    std::vector<int32_t> ident(N);
    std::iota(ident.begin(), ident.end(), 0);
    sbp_bm first = nullptr;
    throw_on(sbp_bm_build(ctx, g, N, ident.data(), &first), ctx, "sbp_bm_build");
    throw_on(sbp_bm_set_model(ctx, first, g, p.nonparametric ? (p.degreecorrected ? 2 : 1) : 0), ctx,
             "sbp_bm_set_model");
    Blockmodel blockmodel = Blockmodel::adopt(ctx, first, N, 1.7976931348623157e308, (long)((double)N * 0.5));
    BlockmodelTriplet triplet;
    uint32_t ep = 0;
    int iteration = 0;
    while (!sbp::done_blockmodeling(blockmodel, triplet)) {
        blockmodel = merge_step(ctx, g, blockmodel, p, ep, st);
        sbp_params q = p;  // finetune::mcmc: async_gibbs only while iteration < asynciterations
        if (p.algorithm == 0 && iteration >= p.asynciterations) q.algorithm = 1;
        mcmc_step(ctx, g, blockmodel, q, triplet.golden_ratio_not_reached(), ep, st);
        ++iteration;
        blockmodel = triplet.get_next_blockmodel(blockmodel);
    }
    if (st) st->mdl = blockmodel.getOverall_entropy();
    return blockmodel;
}

}  // namespace sbp_amd

// ------------------------------------------------------------------ namespaced entry points

namespace entropy {

double mdl(const Blockmodel &bm, const Graph &graph) {
    double m = 0;
    throw_on(sbp_mdl(bm.ctx(), bm.handle(), graph.device(bm.ctx()), &m, nullptr), bm.ctx(), "sbp_mdl");
    return m;
}

double null_mdl_v1(const Graph &graph) {
    if (sbp_amd::args().nonparametric) {  // src/entropy.cpp:572-577: the one-block model's description length
        std::vector<long> zeros(graph.num_vertices(), 0);
        return mdl(Blockmodel(1, graph, 0.5, zeros), graph);
    }
    double E = (double)graph.num_edges();
    double log_posterior_p = E * log(1.0 / E);
    double x = 1.0 / E;
    double h = ((1 + x) * log(1 + x)) - (x * log(x));
    return (E * h) - log_posterior_p;
}

double normalize_mdl_v1(double mdl, const Graph &graph) { return mdl / null_mdl_v1(graph); }

}  // namespace entropy

namespace evaluate {

static sbp_evaluation score_table(long num_vertices, const Hungarian::Matrix &t) {
    if (t.empty() || t[0].empty()) throw std::invalid_argument("evaluate: empty contingency table");
    const int32_t nr = (int32_t)t.size(), nc = (int32_t)t[0].size();
    std::vector<int32_t> flat((size_t)nr * nc);
    for (int32_t r = 0; r < nr; ++r)
        for (int32_t c = 0; c < nc; ++c) flat[(size_t)r * nc + c] = t[r][c];
    sbp_evaluation e{};
    if (sbp_evaluate_table(num_vertices, nr, nc, flat.data(), &e) != SBP_OK)
        throw std::invalid_argument("evaluate: invalid contingency table");
    return e;
}

double calculate_f1_score(long num_vertices, Hungarian::Matrix &contingency_table) {
    const sbp_evaluation e = score_table(num_vertices, contingency_table);
    printf("Accuracy: %g\nRand index: %g\nRecall: %g\nPrecision: %g\nF1 Score: %g\n", e.accuracy, e.rand_index,
           e.recall, e.precision, e.f1_score);
    return e.f1_score;
}

double calculate_nmi(long num_vertices, Hungarian::Matrix &contingency_table) {
    return score_table(num_vertices, contingency_table).nmi;
}

Hungarian::Matrix hungarian(const Graph &graph, Blockmodel &blockmodel) {
    const std::vector<long> &truth = graph.assignment();
    const long N = graph.num_vertices();
    if ((long)truth.size() != N) throw std::invalid_argument("evaluate: the graph has no true assignment");
    std::vector<long> found = blockmodel.block_assignment();
    std::vector<int32_t> t(N), a(N);
    for (long i = 0; i < N; ++i) t[i] = (int32_t)truth[i], a[i] = (int32_t)found[i];
    std::unordered_set<long> labels;
    for (long x : truth)
        if (x >= 0) labels.insert(x);
    long found_blocks = 0;
    for (long x : found) found_blocks = std::max(found_blocks, x + 1);
    std::vector<int32_t> table((size_t)std::max<long>(1, found_blocks * (long)labels.size()));
    int32_t rows = 0, cols = 0;
    sbp_evaluation e{};
    if (sbp_evaluate((int32_t)N, t.data(), a.data(), &e, table.data(), &rows, &cols) != SBP_OK)
        throw std::invalid_argument("evaluate: invalid assignments");
    Hungarian::Matrix m(rows, std::vector<int>(cols));
    for (int32_t r = 0; r < rows; ++r)
        for (int32_t c = 0; c < cols; ++c) m[r][c] = table[(size_t)r * cols + c];
    return m;
}

Eval evaluate_blockmodel(const Graph &graph, Blockmodel &blockmodel) {
    Hungarian::Matrix table = hungarian(graph, blockmodel);
    const double f1 = calculate_f1_score(graph.num_vertices(), table);
    const double nmi = calculate_nmi(graph.num_vertices(), table);
    std::vector<long> truth(graph.assignment());
    const long true_blocks = 1 + *std::max_element(truth.begin(), truth.end());
    Blockmodel true_blockmodel(true_blocks, graph, 0.5, truth);
    return Eval{f1, nmi, entropy::mdl(true_blockmodel, graph)};
}

}  // namespace evaluate

namespace block_merge {

Blockmodel &merge_blocks(Blockmodel &blockmodel, const Graph &graph, long num_edges) {
    (void)num_edges;
    sbp_ctx ctx = blockmodel.ctx();
    blockmodel = merge_step(ctx, graph.device(ctx), blockmodel, to_params(sbp_amd::args()), sbp_amd::epoch(), &g_stats);
    blockmodel.set_graph(&graph);
    return blockmodel;
}

}  // namespace block_merge

namespace divisive {

Blockmodel split_communities(Blockmodel &blockmodel, const Graph &graph, int target_num_communities) {
    sbp_ctx ctx = blockmodel.ctx();
    int32_t num_blocks = 0;
    throw_on(sbp_split_communities(ctx, blockmodel.handle(), graph.device(ctx), (int32_t)target_num_communities,
                                   sbp_amd::epoch()++, &num_blocks, nullptr),
             ctx, "sbp_split_communities");
    blockmodel.set_graph(&graph);
    return blockmodel;
}

}  // namespace divisive

namespace finetune {

static Blockmodel &run_mcmc(Blockmodel &bm, const Graph &graph, int algorithm, bool grnr) {
    sbp_params p = to_params(sbp_amd::args());
    p.algorithm = algorithm;
    mcmc_step(bm.ctx(), graph.device(bm.ctx()), bm, p, grnr, sbp_amd::epoch(), &g_stats);
    return bm;
}

Blockmodel &asynchronous_gibbs(Blockmodel &bm, const Graph &graph, bool grnr) { return run_mcmc(bm, graph, 0, grnr); }
Blockmodel &metropolis_hastings(Blockmodel &bm, const Graph &graph, bool grnr) { return run_mcmc(bm, graph, 1, grnr); }
Blockmodel &hybrid_mcmc(Blockmodel &bm, const Graph &graph, bool grnr) { return run_mcmc(bm, graph, 2, grnr); }

Blockmodel &mcmc(int iteration, const Graph &graph, Blockmodel &bm, BlockmodelTriplet &triplet) {
    const Args &a = sbp_amd::args();
    bool grnr = triplet.golden_ratio_not_reached();
    if ((a.algorithm == "async_gibbs" || a.algorithm == "async_gibbs_load_balanced") && iteration < a.asynciterations)
        return asynchronous_gibbs(bm, graph, grnr);
    if (a.algorithm == "hybrid_mcmc") return hybrid_mcmc(bm, graph, grnr);
    return metropolis_hastings(bm, graph, grnr);
}

Blockmodel &finetune_assignment(Blockmodel &bm, Graph &graph) {
    sbp_params p = to_params(sbp_amd::args());
    double mdl = 0;
    auto t0 = std::chrono::steady_clock::now();
    throw_on(sbp_finetune_assignment(bm.ctx(), bm.handle(), graph.device(bm.ctx()), &p, &sbp_amd::epoch(), &mdl,
                                     &g_stats),
             bm.ctx(), "sbp_finetune_assignment");
    timers::finetune_time += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    bm.setOverall_entropy(mdl);
    bm.set_graph(&graph);
    return bm;
}

}  // namespace finetune

namespace sample {

Sample sample(const Graph &graph) {
    const Args &a = sbp_amd::args();
    static const char *names[] = {"max_degree",         "random", "degree_product", "degree_weighted",
                                  "expansion_snowball", "degree_product_snowball"};
    int32_t alg = -1;
    for (int32_t i = 0; i < 6; ++i)
        if (a.samplingalg == names[i]) alg = i;
    if (alg < 0)
        throw std::invalid_argument(a.samplingalg + " is not a valid sampling algorithm (max_degree, random, "
                                    "degree_product, degree_weighted, expansion_snowball, degree_product_snowball)");
    sbp_ctx ctx = context(a);
    sbp_graph parent = graph.device(ctx);
    sbp_graph s = nullptr;
    int32_t n = 0;
    std::vector<int32_t> mapping(graph.num_vertices());
    throw_on(sbp_sample(ctx, parent, alg, a.samplesize, a.seed, &s, &n, mapping.data()), ctx, "sbp_sample");
    Sample out;
    out.graph = Graph::adopt(ctx, s);
    out.mapping.assign(mapping.begin(), mapping.end());
    std::vector<long> truth(n, -1);  // from_vertices carries the true assignment over (sample.cpp:214)
    const std::vector<long> &full_truth = graph.assignment();
    if (!full_truth.empty())
        for (long v = 0; v < graph.num_vertices(); ++v)
            if (out.mapping[v] >= 0) truth[out.mapping[v]] = full_truth[v];
    out.graph.assignment(std::move(truth));
    return out;
}

std::vector<long> extend(const Graph &graph, const Blockmodel &sample_blockmodel, const Sample &sample) {
    sbp_ctx ctx = sample_blockmodel.ctx();
    std::vector<int32_t> a(graph.num_vertices());
    sbp_bm full = nullptr;
    throw_on(sbp_extend(ctx, graph.device(ctx), sample.graph.device(ctx), sample_blockmodel.handle(),
                        sbp_amd::args().seed, &full, a.data()),
             ctx, "sbp_extend");
    sbp_bm_free(full);
    return std::vector<long>(a.begin(), a.end());
}

Sample detach(const Graph &graph) {
    sbp_ctx ctx = context(sbp_amd::args());
    sbp_graph d = nullptr;
    int32_t n = 0;
    std::vector<int32_t> mapping(graph.num_vertices());
    throw_on(sbp_detach(ctx, graph.device(ctx), &d, &n, mapping.data()), ctx, "sbp_detach");
    Sample out;
    out.graph = Graph::adopt(ctx, d);
    out.mapping.assign(mapping.begin(), mapping.end());
    std::vector<long> truth(n, -1);
    const std::vector<long> &full_truth = graph.assignment();
    if (!full_truth.empty())
        for (long v = 0; v < graph.num_vertices(); ++v)
            if (out.mapping[v] >= 0) truth[out.mapping[v]] = full_truth[v];
    out.graph.assignment(std::move(truth));
    return out;
}

Blockmodel reattach(const Graph &graph, const Blockmodel &sample_blockmodel, const Sample &sample) {
    const long N = graph.num_vertices();
    std::vector<int32_t> mapping(sample.mapping.begin(), sample.mapping.end());
    std::vector<long> sa = sample_blockmodel.block_assignment();
    std::vector<int32_t> detached(sa.begin(), sa.end()), assignment(N);
    if (sbp_reattach((int32_t)N, graph.out_off().data(), graph.out_adj().data(), graph.in_off().data(),
                     graph.in_adj().data(), mapping.data(), detached.data(), (int32_t)sample_blockmodel.num_blocks(),
                     sbp_amd::args().seed, assignment.data()) != SBP_OK)
        throw std::invalid_argument("sample::reattach: invalid arguments");
    std::vector<long> a(assignment.begin(), assignment.end());
    return Blockmodel(sample_blockmodel.num_blocks(), graph, 0.5, a);
}

}  // namespace sample

namespace sbp {

bool done_blockmodeling(Blockmodel &blockmodel, BlockmodelTriplet &triplet, long min_num_blocks) {
    if (min_num_blocks > 0) {
        if (blockmodel.num_blocks() <= min_num_blocks || !triplet.get(2).empty) return true;
    }
    return triplet.optimal_num_blocks_found;
}

const sbp_stats &last_stats() { return g_stats; }

static long g_last_num_blocks = 0;
long last_num_blocks() { return g_last_num_blocks; }

Blockmodel stochastic_block_partition(Graph &graph, Args &a, bool divide_and_conquer) {
    (void)divide_and_conquer;  // divide-and-conquer is an MPI mode, not part of this path
    if (&a != &::args) sbp_amd::args() = a;
    sbp_amd::epoch() = 0;
    memset(&g_stats, 0, sizeof(g_stats));
    auto t0 = std::chrono::steady_clock::now();
    sbp_ctx ctx = context(a);
    Blockmodel result = drive(ctx, graph.device(ctx), (int32_t)graph.num_vertices(), to_params(a), &g_stats);
    g_stats.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    result.set_graph(&graph);
    g_last_num_blocks = result.num_blocks();
    // the reference's phase timers and counters (include/globals.hpp:19-67), in seconds
    timers::MCMC_moves += (unsigned long)g_stats.moves;
    timers::MCMC_iterations += g_stats.sweeps;
    timers::MCMC_time += g_stats.finetune_ms / 1e3;
    timers::MCMC_parallel_time += g_stats.ft_kernel_ms / 1e3;
    timers::BlockMerge_time += (g_stats.merge_ms + g_stats.carry_ms) / 1e3;
    timers::BlockMerge_loop_time += g_stats.merge_ms / 1e3;
    timers::BLOCKMODEL_BUILD_TIME += g_stats.build_ms / 1e3;
    timers::total_time += g_stats.total_ms / 1e3;
    return result;
}

}  // namespace sbp
