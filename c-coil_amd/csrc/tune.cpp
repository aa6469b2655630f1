// The autotuner: per call size, the launch shape that measured fastest on this device, kept in the context's cache
// and optionally in a text file, so that later calls (and later processes) replay the same choice.
#include "context.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <set>
#include <sstream>

using namespace ccoil;

namespace ccoil {

namespace {

bool cache_get(ccoil_ctx ctx, const std::string &key, std::string &value)
{
    auto it = ctx->tuning.find(key);
    if (it == ctx->tuning.end()) return false;
    value = it->second;
    return true;
}

void cache_put(ccoil_ctx ctx, const std::string &key, const std::string &value)
{
    ctx->tuning[key] = value;
    if (ctx->tuning_file.empty()) return;
    std::ofstream f(ctx->tuning_file, std::ios::app);
    if (f) f << key << '\t' << value << '\n';
}

}  // namespace

ccoil_status load_tuning_file(ccoil_ctx ctx, const char *path)
{
    ctx->tuning_file = path ? path : "";
    if (!path) return CCOIL_OK;
    std::ifstream f(path);
    std::string line;
    while (f && std::getline(f, line)) {
        const size_t tab = line.find('\t');
        if (tab == std::string::npos || tab == 0) continue;
        ctx->tuning[line.substr(0, tab)] = line.substr(tab + 1);
    }
    return CCOIL_OK;
}

ccoil_status tune_fields(ccoil_ctx ctx, const ccoil_coils *coils, const double *points, int64_t n_points,
                         uint32_t quantities, int32_t precision, uint32_t flags, double *A, double *B, double *E,
                         double *G, FieldShape &out)
{
    std::vector<CoilDev> hc;
    const std::string msg = pack_coils(coils, false, (quantities & CCOIL_E) != 0, hc);
    if (!msg.empty() || n_points <= 0 || hc.empty()) {   // fields_call reports it, or has nothing to tune
        out = FieldShape();
        return CCOIL_OK;
    }
    int64_t nodes = 0;
    for (const auto &c : hc) nodes += (int64_t)c.n_ang * c.n_thk;
    std::ostringstream k;
    k << "fields n=" << n_points << " coils=" << hc.size() << " nodes=" << nodes << " q=" << quantities
      << " precision=" << precision << " near=" << ((flags & CCOIL_NO_NEAR_FP64) ? 0 : 1) << " cu=" << ctx->cu_count;
    const std::string key = k.str();
    std::string v;
    if (cache_get(ctx, key, v) && std::sscanf(v.c_str(), "split=%d chunks=%d", &out.split, &out.chunks) == 2)
        return CCOIL_OK;
    // candidates around the rule: every lane count that can fill the chip, 1-4 coil chunks and the rule's chunks
    // with its halves and doubles
    const int64_t nc = (int64_t)hc.size();
    const FieldShape rule = field_shape_rule(ctx, n_points, nc, quantities, precision);
    std::set<std::pair<int, int>> cand;
    std::vector<int> splits = {1, 4};
    if (n_points < 65536) splits.push_back(16);
    for (int s : splits)
        for (int64_t c : {(int64_t)1, (int64_t)2, (int64_t)3, (int64_t)4, (int64_t)rule.chunks,
                          (int64_t)std::max(1, rule.chunks / 2), (int64_t)rule.chunks * 2})
            if (c >= 1 && c <= nc && c <= 65535) cand.insert({s, (int)c});
    cand.insert({rule.split, rule.chunks});
    double best = 1e300;
    FieldShape best_shape = rule;
    for (const auto &c : cand) {
        FieldShape sh;
        sh.split = c.first;
        sh.chunks = c.second;
        ccoil_report rep;
        const ccoil_status st = fields_call(ctx, coils, points, n_points, quantities, precision, flags, A, B, E, G,
                                            &rep, nullptr, &sh);
        if (st != CCOIL_OK) return st;
        if (rep.main_kernel_ms < best) {
            best = rep.main_kernel_ms;
            best_shape = sh;
        }
    }
    out = best_shape;
    std::ostringstream o;
    o << "split=" << out.split << " chunks=" << out.chunks;
    cache_put(ctx, key, o.str());
    return CCOIL_OK;
}

ccoil_status tune_pairs(ccoil_ctx ctx, const PairBuild &rule, uint32_t what, int32_t precision,
                        const std::function<ccoil_status(int64_t, double &)> &run, int64_t &target)
{
    std::ostringstream k;
    k << "pairs n=" << rule.sorted.size() << " node_pairs=" << rule.node_pairs << " what=" << what
      << " precision=" << precision << " cu=" << ctx->cu_count;
    const std::string key = k.str();
    std::string v;
    long long t = 0;
    if (cache_get(ctx, key, v) && std::sscanf(v.c_str(), "target=%lld", &t) == 1 && t > 0) {
        target = t;
        return CCOIL_OK;
    }
    double best = 1e300;
    target = rule.target;
    for (int64_t c : {rule.target / 4, rule.target / 2, rule.target, rule.target * 2}) {
        if (c < 256 || c > 32768) continue;
        double ms = 0.0;
        const ccoil_status st = run(c, ms);
        if (st != CCOIL_OK) return st;
        if (ms < best) {
            best = ms;
            target = c;
        }
    }
    cache_put(ctx, key, "target=" + std::to_string(target));
    return CCOIL_OK;
}

}  // namespace ccoil

extern "C" ccoil_status ccoil_set_tuning_file(ccoil_ctx ctx, const char *path)
{
    if (!ctx) return fail(nullptr, CCOIL_ERR_ARGUMENT, "ccoil_set_tuning_file: ctx is NULL");
    return load_tuning_file(ctx, path);
}
