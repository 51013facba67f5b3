// esl_compat.hpp -- the reference's C++ class API (cuda/src/*.h of
// asmelko/paraExaStoLog) over the C ABI of libesl.so, header-only.
// Members are host std::vectors instead of thrust::device_vectors; errors
// are std::runtime_error, as the reference's input errors.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>
#include "esl.h"
#include "esl_symb.hpp"
#include <algorithm>
#include <cstdint>
#include <iostream>
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
struct transition_table;
// the tables alive in this thread: transition_graph(ctx, rows, cols, indptr)
// finds the table that owns the vectors it is given, as the reference keeps
// const references to them (transition_graph.h:10)
inline std::vector<transition_table*>& live_tables() { static thread_local std::vector<transition_table*> v; return v; }
struct transition_table {                 // transition_table.h:7-26
    cu_context& ctx; const model_t& model; std::vector<double> r; esl_graph* g{};
    std::vector<int32_t> indptr, rows, cols;
    transition_table(cu_context& c, const model_t& m, const std::vector<double>& rates) : ctx(c), model(m), r(rates) {
        live_tables().push_back(this);
    }
    void construct_table() {
        ok(esl_table_build(ctx.h, model.h, r.data(), &g), ctx.h);
        int64_t s[7]; ok(esl_graph_query(g, s));
        indptr.resize(s[0] + 1); rows.resize(s[1]); cols.resize(s[1]);
        ok(esl_table_export(ctx.h, g, indptr.data(), rows.data(), cols.data()), ctx.h);
    }
    transition_table(const transition_table&) = delete;
    transition_table& operator=(const transition_table&) = delete;
    ~transition_table() {
        auto& v = live_tables();
        v.erase(std::remove(v.begin(), v.end(), this), v.end());
        esl_graph_destroy(g);
    }
};
struct transition_graph {                 // transition_graph.h:6-28
    cu_context& ctx; transition_table& t;
    int32_t terminals_count{}; std::vector<int32_t> sccs_offsets, reordered_vertices;
    transition_graph(cu_context& c, transition_table& tab) : ctx(c), t(tab) {}
    // the reference's signature (transition_graph.h:21): the table's own vectors
    transition_graph(cu_context& c, const std::vector<int32_t>& rows, const std::vector<int32_t>& cols,
                     const std::vector<int32_t>& indptr) : ctx(c), t(owner(rows, cols, indptr)) {}
    void reorganize_vertices() {
        ok(esl_graph_reorganize(ctx.h, t.g), ctx.h);
        int64_t s[7]; ok(esl_graph_query(t.g, s));
        reordered_vertices.resize(s[0]); sccs_offsets.resize(s[2] + 1);
        ok(esl_graph_export(ctx.h, t.g, reordered_vertices.data(), sccs_offsets.data(), &terminals_count, nullptr), ctx.h);
    }
private:
    static transition_table& owner(const std::vector<int32_t>& rows, const std::vector<int32_t>& cols,
                                   const std::vector<int32_t>& indptr) {
        for (transition_table* tab : live_tables())
            if (&tab->rows == &rows && &tab->cols == &cols && &tab->indptr == &indptr) return *tab;
        throw std::runtime_error("transition_graph: rows, cols and indptr must be the members of a live transition_table");
    }
};
template <bool CSR>
struct sparse_cs_matrix {                 // sparse_utils.h:12-50 (host vectors, FP64 values)
    std::vector<int32_t> indptr, indices;
    std::vector<double> data;
    int64_t nnz() const { return (int64_t)indices.size(); }
    int64_t n() const { return indptr.empty() ? 0 : (int64_t)indptr.size() - 1; }
};
using sparse_csc_matrix = sparse_cs_matrix<false>;
using sparse_csr_matrix = sparse_cs_matrix<true>;
using persistent_data = esl_symb::data_t; // persistent_solution.h:5-24 (the .symb file's arrays)

struct solver {                           // solver.h:10-67
    cu_context& ctx; transition_table* t{}; const transition_graph* gr{};
    const model_t& model; std::vector<double> rates, x0;
    persistent_data* persisted{};
    sparse_csc_matrix solution_term;      // n x T: terminal SCC j's kernel in column j
    sparse_csr_matrix solution_nonterm;   // T x n: absorption probabilities
    std::vector<double> final_state;
    bool recompute_needed = false;
    solver(cu_context& c, transition_table& tab, const transition_graph& g, const transition_rates& r, const initial_state& s)
        : ctx(c), t(&tab), gr(&g), model(r.model), rates(r.rates), x0(s.state), final_state(s.state.size()) {}
    // solver(ctx, persisted, rates, state) (solver.h:49, solver.cu:25-51): with the
    // stored rates the stored solution is reused; otherwise solve() recomputes on
    // the stored table and graph stage (same zero pattern required)
    solver(cu_context& c, persistent_data& p, const transition_rates& r, const initial_state& s)
        : ctx(c), model(r.model), rates(r.rates), x0(s.state), persisted(&p), final_state(s.state.size()) {
        recompute_needed = !esl_symb::same_rates(p, rates);
        if (!recompute_needed) {
            solution_term = {p.term_indptr, p.term_indices, std::vector<double>(p.term_data.begin(), p.term_data.end())};
            solution_nonterm = {p.nonterm_indptr, p.nonterm_indices,
                                std::vector<double>(p.nonterm_data.begin(), p.nonterm_data.end())};
        }
    }
    ~solver() { if (own) esl_graph_destroy(own); }
    solver(const solver&) = delete;
    solver& operator=(const solver&) = delete;
    void solve() {
        if (persisted && !recompute_needed) {       // R (L x0) from the stored solution (solver.cu:302-309)
            final_state = esl_symb::final_from_solution(*persisted, x0);
            return;
        }
        esl_graph* g = t ? t->g : nullptr;
        if (persisted) {                            // stored graph stage, new rates of the same zero pattern
            if (!esl_symb::compatible_zero_rates(*persisted, rates))
                throw std::runtime_error("rates are not compatible with the stored symbolic data");
            if (persisted->ordered_vertices.size() != x0.size() || persisted->terminals_offsets.empty() ||
                persisted->nonterminals_offsets.empty())
                throw std::runtime_error("persistent data file is corrupted");
            ok(esl_table_build(ctx.h, model.h, rates.data(), &own), ctx.h);
            std::vector<int32_t> offs(persisted->terminals_offsets);
            offs.insert(offs.end(), persisted->nonterminals_offsets.begin() + 1, persisted->nonterminals_offsets.end());
            ok(esl_graph_import(ctx.h, own, persisted->ordered_vertices.data(), offs.data(), (int64_t)offs.size() - 1,
                                (int32_t)persisted->terminals_offsets.size() - 1), ctx.h);
            g = own;
        }
        esl_solve_opts o; esl_solve_opts_default(&o);
        o.full = 1;                                 // solution_term / solution_nonterm, as solve() always builds them
        ok(esl_solve(ctx.h, g, rates.data(), x0.data(), final_state.data(), &o, nullptr), ctx.h);
        int64_t sz[7]; ok(esl_graph_query(g, sz));
        const int64_t T = sz[3], n_t = sz[4];
        int64_t nnz_l = 0;
        ok(esl_solution_export(ctx.h, g, nullptr, nullptr, nullptr, &nnz_l, nullptr, nullptr, nullptr), ctx.h);
        solution_term = {std::vector<int32_t>(T + 1), std::vector<int32_t>(n_t), std::vector<double>(n_t)};
        solution_nonterm = {std::vector<int32_t>(T + 1), std::vector<int32_t>(nnz_l), std::vector<double>(nnz_l)};
        ok(esl_solution_export(ctx.h, g, solution_term.indptr.data(), solution_term.indices.data(),
                               solution_term.data.data(), &nnz_l, solution_nonterm.indptr.data(),
                               solution_nonterm.indices.data(), solution_nonterm.data.data()), ctx.h);
    }
    // solver.cu:336-372: non-zero states (as stored, FP32) as their nodes joined by " --- ",
    // "<nil>" for state 0, then the value with the stream's default precision
    void print_final_state(const std::vector<std::string>& model_nodes, std::ostream& os = std::cout) const {
        for (size_t s = 0; s < final_state.size(); ++s) {
            const float v = (float)final_state[s];
            if (v == 0.0f) continue;
            bool any = false;
            for (size_t i = 0; i < model_nodes.size(); ++i)
                if (s & (size_t(1) << i)) { os << (any ? " --- " : "") << model_nodes[i]; any = true; }
            if (!any) os << "<nil>";
            os << " " << v << std::endl;
        }
    }
    friend struct persistent_solution;
private:
    esl_graph* own{};                     // the table rebuilt for a persisted solve
};

struct persistent_solution {              // persistent_solution.h:27-37 (.symb, esl_symb.hpp)
    static persistent_data deserialize(const std::string& file, bool no_inverse) {
        return esl_symb::read(file, !no_inverse);
    }
    // the table, graph stage, rates and solution of a solved solver; no inverse is
    // written (this library factors nothing), as with no_inverse = true
    static void serialize(const std::string& file, const solver& s, bool /*no_inverse*/) {
        if (!s.t || !s.gr) throw std::runtime_error("serialize: the solver was not built from a table and a graph");
        persistent_data d;
        d.rows = s.t->rows; d.cols = s.t->cols; d.indptr = s.t->indptr;
        d.ordered_vertices = s.gr->reordered_vertices;
        const size_t T = (size_t)s.gr->terminals_count;
        d.terminals_offsets.assign(s.gr->sccs_offsets.begin(), s.gr->sccs_offsets.begin() + T + 1);
        d.nonterminals_offsets.assign(s.gr->sccs_offsets.begin() + T, s.gr->sccs_offsets.end());
        for (double v : s.rates) d.rates.push_back((float)v);
        d.term_indptr = s.solution_term.indptr; d.term_indices = s.solution_term.indices;
        for (double v : s.solution_term.data) d.term_data.push_back((float)v);
        d.nonterm_indptr = s.solution_nonterm.indptr; d.nonterm_indices = s.solution_nonterm.indices;
        for (double v : s.solution_nonterm.data) d.nonterm_data.push_back((float)v);
        esl_symb::write(file, d);
    }
    static bool has_compatible_zero_rates(const persistent_data& d, const std::vector<double>& r) {
        return esl_symb::compatible_zero_rates(d, r);
    }
    static bool are_same(const persistent_data& d, const std::vector<double>& r) { return esl_symb::same_rates(d, r); }
    // check_are_equal (persistent_solution.cu:247-269): every stored array equals
    // the solver's, values compared as the file holds them (FP32)
    static bool check_are_equal(const solver& s, const persistent_data& d) {
        if (!s.t || !s.gr) throw std::runtime_error("check_are_equal: the solver was not built from a table and a graph");
        const auto same = [](const std::vector<int32_t>& a, const std::vector<int32_t>& b) { return a == b; };
        const auto same_f = [](const std::vector<double>& a, const std::vector<float>& b) {
            if (a.size() != b.size()) return false;
            for (size_t i = 0; i < a.size(); ++i)
                if ((float)a[i] != b[i]) return false;
            return true;
        };
        const auto& off = s.gr->sccs_offsets;
        const size_t T = (size_t)s.gr->terminals_count;
        if (off.size() < T + 1) return false;
        bool eq = same(s.t->rows, d.rows) && same(s.t->cols, d.cols) && same(s.t->indptr, d.indptr);
        eq = eq && same(s.gr->reordered_vertices, d.ordered_vertices);
        eq = eq && same(std::vector<int32_t>(off.begin(), off.begin() + T + 1), d.terminals_offsets);
        eq = eq && same(std::vector<int32_t>(off.begin() + T, off.end()), d.nonterminals_offsets);
        eq = eq && same_f(s.rates, d.rates);
        eq = eq && same(s.solution_term.indptr, d.term_indptr) && same(s.solution_term.indices, d.term_indices);
        eq = eq && same_f(s.solution_term.data, d.term_data);
        eq = eq && same(s.solution_nonterm.indptr, d.nonterm_indptr) && same(s.solution_nonterm.indices, d.nonterm_indices);
        return eq && same_f(s.solution_nonterm.data, d.nonterm_data);
    }
};
} // namespace compat
