// esl_compat.hpp -- the reference's C++ class API (cuda/src/*.h of
// asmelko/paraExaStoLog) over the C ABI of libesl.so, header-only.
// Members are host std::vectors instead of thrust::device_vectors; errors
// are std::runtime_error, as the reference's input errors.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>
#include "esl.h"
#include <cstdint>
#include <memory>
#include <utility>

namespace compat {
inline void ok(esl_status s, const esl_ctx* c = nullptr) {
    if (s != ESL_OK) throw std::runtime_error(esl_last_error(c)); // reference: runtime_error / exit
}
struct cu_context {                       // replaces cu_context (cuda/src/cu_context.h:15-23)
    esl_ctx* h{};
    cu_context() { ok(esl_ctx_create(0, &h)); }
    cu_context(const cu_context&) = delete;
    cu_context& operator=(const cu_context&) = delete;
    ~cu_context() { esl_ctx_destroy(h); }
};
struct model_t {                          // model.h:28-32
    esl_model* h{};
    std::shared_ptr<esl_model> owner;     // releases the handle with the last copy
    std::vector<std::string> nodes;
};
struct model_builder {                    // model.h:34-42
    model_t construct_model(const std::string& f) {
        model_t m; ok(esl_model_load(f.c_str(), &m.h));
        m.owner.reset(m.h, esl_model_destroy);
        for (int i = 0; i < esl_model_num_nodes(m.h); ++i) m.nodes.push_back(esl_model_node_name(m.h, i));
        return m;
    }
};
using ptrans_t = std::pair<std::string, double>;
struct transition_rates {                 // transition_rates.h:12-31
    const model_t& model; std::vector<double> rates;
    explicit transition_rates(const model_t& m) : model(m), rates(2 * m.nodes.size()) {}
    void fill(int kind, double mean, double sd, const std::vector<ptrans_t>& up, const std::vector<ptrans_t>& dn) {
        std::vector<const char*> un, dnn; std::vector<double> uv, dv;
        for (auto& p : up) { un.push_back(p.first.c_str()); uv.push_back(p.second); }
        for (auto& p : dn) { dnn.push_back(p.first.c_str()); dv.push_back(p.second); }
        ok(esl_rates_fill(model.h, kind, mean, sd, un.data(), uv.data(), (int)un.size(),
                          dnn.data(), dv.data(), (int)dnn.size(), rates.data()));
    }
    void generate_uniform(const std::vector<ptrans_t>& up = {}, const std::vector<ptrans_t>& dn = {}) { fill(0, 0, 0, up, dn); }
    void generate_normal(double mean, double sd, const std::vector<ptrans_t>& up = {}, const std::vector<ptrans_t>& dn = {}) { fill(1, mean, sd, up, dn); }
};
struct initial_state {                    // initial_state.h:5-11
    std::vector<double> state;
    initial_state(const model_t& m, const std::vector<std::string>& fixed = {},
                  const std::vector<bool>& vals = {}, double p = 1.0) : state(size_t(1) << m.nodes.size()) {
        std::vector<const char*> f; std::vector<uint8_t> v;
        for (auto& s : fixed) f.push_back(s.c_str());
        for (bool b : vals) v.push_back(b);
        if (f.size() != v.size()) throw std::runtime_error("number of fixed node names not equal to number of fixed node values");
        ok(esl_initial_state(m.h, f.data(), v.data(), (int)f.size(), p, state.data()));
    }
};
struct transition_table {                 // transition_table.h:7-26
    cu_context& ctx; const model_t& model; std::vector<double> r; esl_graph* g{};
    std::vector<int32_t> indptr, rows, cols;
    transition_table(cu_context& c, const model_t& m, const std::vector<double>& rates) : ctx(c), model(m), r(rates) {}
    void construct_table() {
        ok(esl_table_build(ctx.h, model.h, r.data(), &g), ctx.h);
        int64_t s[7]; ok(esl_graph_query(g, s));
        indptr.resize(s[0] + 1); rows.resize(s[1]); cols.resize(s[1]);
        ok(esl_table_export(ctx.h, g, indptr.data(), rows.data(), cols.data()), ctx.h);
    }
    transition_table(const transition_table&) = delete;
    transition_table& operator=(const transition_table&) = delete;
    ~transition_table() { esl_graph_destroy(g); }
};
struct transition_graph {                 // transition_graph.h:6-28
    cu_context& ctx; transition_table& t;
    int32_t terminals_count{}; std::vector<int32_t> sccs_offsets, reordered_vertices;
    transition_graph(cu_context& c, transition_table& tab) : ctx(c), t(tab) {}
    void reorganize_vertices() {
        ok(esl_graph_reorganize(ctx.h, t.g), ctx.h);
        int64_t s[7]; ok(esl_graph_query(t.g, s));
        reordered_vertices.resize(s[0]); sccs_offsets.resize(s[2] + 1);
        ok(esl_graph_export(ctx.h, t.g, reordered_vertices.data(), sccs_offsets.data(), &terminals_count, nullptr), ctx.h);
    }
};
struct solver {                           // solver.h:10-67
    cu_context& ctx; transition_table& t; std::vector<double> rates, x0, final_state;
    solver(cu_context& c, transition_table& tab, const transition_graph&, const transition_rates& r, const initial_state& s)
        : ctx(c), t(tab), rates(r.rates), x0(s.state), final_state(s.state.size()) {}
    void solve() {
        esl_solve_opts o; esl_solve_opts_default(&o);
        ok(esl_solve(ctx.h, t.g, rates.data(), x0.data(), final_state.data(), &o, nullptr), ctx.h);
    }
};
} // namespace compat
