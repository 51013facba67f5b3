// C ABI of libesl.so (declared in include/esl.h).  Every entry point catches
// esl::error / std::exception and returns a status; nothing is thrown across
// the boundary and nothing calls exit() (the reference exits the process on a
// device error, cuda/src/cu_context.cpp:7-32).

#include <algorithm>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <random>
#include <sstream>

#include "esl_internal.hpp"

struct esl_ctx
{
	esl::ctx_t c;
};

struct esl_model
{
	esl::model_t m;
};

struct esl_graph
{
	esl::graph_t g;
};

struct esl_plan
{
	esl_ctx* ctx = nullptr;
	const esl_graph* graph = nullptr;
	void* p = nullptr;
	~esl_plan()
	{
		if (p)
			esl::plan_destroy(p);
	}
};

namespace {

thread_local std::string tl_error;

template <typename F>
esl_status guarded(esl_ctx* ctx, F&& f)
{
	try
	{
		f();
		return ESL_OK;
	}
	catch (const esl::error& e)
	{
		(ctx ? ctx->c.last_error : tl_error) = e.what();
		return (esl_status)e.code;
	}
	catch (const std::bad_alloc&)
	{
		(ctx ? ctx->c.last_error : tl_error) = "host allocation failed";
		return ESL_E_OOM;
	}
	catch (const std::exception& e)
	{
		(ctx ? ctx->c.last_error : tl_error) = e.what();
		return ESL_E_ARG;
	}
}

void require(bool cond, int code, const std::string& what)
{
	if (!cond)
		throw esl::error(code, what);
}

int node_index(const esl::model_t& m, const char* name)
{
	require(name != nullptr, ESL_E_ARG, "null node name");
	auto it = std::find(m.nodes.begin(), m.nodes.end(), std::string(name));
	if (it == m.nodes.end())
		return -1;
	return (int)(it - m.nodes.begin());
}

void set_device(esl_ctx* ctx) { ESL_HIP(hipSetDevice(ctx->c.device)); }

void check_rates(const esl::graph_t& g, const double* rates)
{
	for (int i = 0; i < 2 * g.N; ++i)
	{
		require(rates[i] >= 0.0, ESL_E_ARG, "transition rate must be non-negative");
		require((rates[i] == 0.0) == (g.rates[i] == 0.0), ESL_E_ARG,
				"rates have a different zero pattern than the table was built with");
	}
}

} // namespace

esl::ctx_t& esl::ctx_impl(esl_ctx* ctx) { return ctx->c; }

extern "C" {

const char* esl_version(void) { return "paraexastolog-amd 0.1 (gfx950)"; }

esl_status esl_ctx_create(int device, esl_ctx** out)
{
	return guarded(nullptr, [&] {
		require(out != nullptr, ESL_E_ARG, "null output");
		int count = 0;
		if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
			throw esl::error(ESL_E_DEVICE, "no HIP device available");
		require(device >= 0 && device < count, ESL_E_ARG, "device index out of range");
		auto ctx = std::make_unique<esl_ctx>();
		ctx->c.device = device;
		ESL_HIP(hipSetDevice(device));
		ESL_HIP(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking));
		int lds = 0;
		ESL_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
		if (lds > 0)
			ctx->c.lds_per_block = lds;
		for (auto& e : ctx->c.ev)
			ESL_HIP(hipEventCreate(&e));
		ctx->c.flags.ensure(64);
		ESL_HIP(hipHostMalloc((void**)&ctx->c.h_flags, 64 * sizeof(int32_t), hipHostMallocDefault));
		ctx->c.h_stage_bytes = esl::stage_bytes;
		ESL_HIP(hipHostMalloc((void**)&ctx->c.h_stage, ctx->c.h_stage_bytes, hipHostMallocDefault));
		esl::noalloc_init(ctx->c);
		int wclk = 0;
		if (hipDeviceGetAttribute(&wclk, hipDeviceAttributeWallClockRate, device) == hipSuccess && wclk > 0)
			ctx->c.wall_khz = wclk;
		*out = ctx.release();
	});
}

void esl_ctx_destroy(esl_ctx* ctx)
{
	if (!ctx)
		return;
	(void)hipSetDevice(ctx->c.device);
	if (ctx->c.stream)
		(void)hipStreamSynchronize(ctx->c.stream);
	for (auto& e : ctx->c.ev)
		if (e)
			(void)hipEventDestroy(e);
	if (ctx->c.h_flags)
		(void)hipHostFree(ctx->c.h_flags);
	if (ctx->c.h_stage)
		(void)hipHostFree(ctx->c.h_stage);
	ctx->c.flags.reset();
	ctx->c.scratch_a.reset();
	ctx->c.scratch_b.reset();
	if (ctx->c.multi && ctx->c.multi_free)
		ctx->c.multi_free(ctx->c.multi);
	for (auto s : ctx->c.aux)
	{
		(void)hipStreamSynchronize(s);
		(void)hipStreamDestroy(s);
	}
	for (auto e : ctx->c.aux_ev)
		(void)hipEventDestroy(e);
	if (ctx->c.fill_s)
	{
		(void)hipStreamSynchronize(ctx->c.fill_s);
		(void)hipStreamDestroy(ctx->c.fill_s);
	}
	for (auto e : ctx->c.fill_ev)
		if (e)
			(void)hipEventDestroy(e);
	if (ctx->c.stream)
		(void)hipStreamDestroy(ctx->c.stream);
	delete ctx;
}

const char* esl_last_error(const esl_ctx* ctx) { return ctx ? ctx->c.last_error.c_str() : tl_error.c_str(); }

// ---------------------------------------------------------------- model

esl_status esl_model_parse(const char* text, esl_model** out)
{
	return guarded(nullptr, [&] {
		require(text && out, ESL_E_ARG, "null argument");
		auto m = std::make_unique<esl_model>();
		m->m = esl::parse_bnet(text);
		*out = m.release();
	});
}

esl_status esl_model_load(const char* path, esl_model** out)
{
	return guarded(nullptr, [&] {
		require(path && out, ESL_E_ARG, "null argument");
		std::ifstream f(path, std::ios::binary);
		if (!f.is_open())
			throw esl::error(ESL_E_IO, std::string("could not open model file ") + path);
		std::stringstream ss;
		ss << f.rdbuf();
		auto m = std::make_unique<esl_model>();
		m->m = esl::parse_bnet(ss.str());
		*out = m.release();
	});
}

void esl_model_destroy(esl_model* m) { delete m; }

int esl_model_num_nodes(const esl_model* m) { return m ? (int)m->m.nodes.size() : 0; }

const char* esl_model_node_name(const esl_model* m, int i)
{
	if (!m || i < 0 || i >= (int)m->m.nodes.size())
		return nullptr;
	return m->m.nodes[i].c_str();
}

esl_status esl_model_rule(const esl_model* m, int i, int* k_out, int32_t* inputs, uint8_t* table_bits)
{
	return guarded(nullptr, [&] {
		require(m && k_out, ESL_E_ARG, "null argument");
		require(i >= 0 && i < (int)m->m.nodes.size(), ESL_E_ARG, "node index out of range");
		const auto& r = m->m.rules[i];
		*k_out = (int)r.inputs.size();
		if (inputs)
			std::copy(r.inputs.begin(), r.inputs.end(), inputs);
		if (table_bits)
			std::copy(r.table.begin(), r.table.end(), table_bits);
	});
}

// ---------------------------------------------------------------- rates / x0

esl_status esl_rates_fill(const esl_model* m, int kind, double mean, double std_dev, const char* const* up_names,
						  const double* up_vals, int n_up, const char* const* down_names, const double* down_vals,
						  int n_down, double* rates_out)
{
	return guarded(nullptr, [&] {
		require(m && rates_out, ESL_E_ARG, "null argument");
		require(kind == 0 || kind == 1, ESL_E_ARG, "rate kind must be 0 (UNIFORM) or 1 (NORMAL)");
		const int N = (int)m->m.nodes.size();
		// transition_rates::transform (transition_rates.cpp:7-22)
		auto transform = [&](const char* const* names, const double* vals, int cnt) {
			std::map<int, double> r;
			for (int k = 0; k < cnt; ++k)
			{
				const int i = node_index(m->m, names[k]);
				if (i < 0)
					throw esl::error(ESL_E_ARG, std::string("node ") + names[k] + " not found");
				if (vals[k] < 0)
					throw esl::error(ESL_E_ARG, "transition rate must be non-negative");
				r[i] = vals[k];
			}
			return r;
		};
		auto up = transform(up_names, up_vals, n_up);
		auto down = transform(down_names, down_vals, n_down);
		// generator: transition_rates.cpp:62-97 (float draws, as the reference's real_t)
		std::default_random_engine engine;
		std::normal_distribution<float> dist((float)mean, (float)std_dev);
		auto draw = [&]() -> double {
			if (kind == 0)
				return 1.0;
			float x;
			size_t count = 0;
			do
			{
				x = dist(engine);
				++count;
			} while (x < 0 && count < 1000);
			if (x < 0)
				throw esl::error(ESL_E_ARG,
								 "could not generate postive transition rates with the provided mean and std");
			return (double)x;
		};
		for (int i = 0; i < N; ++i)
		{
			auto u = up.find(i);
			rates_out[2 * i] = u != up.end() ? u->second : draw();
			auto d = down.find(i);
			rates_out[2 * i + 1] = d != down.end() ? d->second : draw();
		}
	});
}

esl_status esl_initial_state(const esl_model* m, const char* const* fixed_names, const uint8_t* fixed_vals,
							 int n_fixed, double p, double* x0_out)
{
	return guarded(nullptr, [&] {
		require(m && x0_out, ESL_E_ARG, "null argument");
		const int N = (int)m->m.nodes.size();
		const uint64_t n = 1ull << N;
		if (n_fixed <= 0)
		{
			std::fill(x0_out, x0_out + n, 1.0 / (double)n);
			return;
		}
		require(fixed_names && fixed_vals, ESL_E_ARG, "null fixed-node arrays");
		uint32_t fmask = 0, fval = 0;
		for (int k = 0; k < n_fixed; ++k)
		{
			const int i = node_index(m->m, fixed_names[k]);
			if (i < 0)
				throw esl::error(ESL_E_ARG, std::string("invalid initial node name: ") + fixed_names[k]);
			if (fmask & (1u << i))
				throw esl::error(ESL_E_ARG, std::string("node fixed twice: ") + fixed_names[k]);
			fmask |= 1u << i;
			if (fixed_vals[k])
				fval |= 1u << i;
		}
		const uint64_t nf = n >> __builtin_popcount(fmask);
		const double pin = p / (double)nf;
		const double pout = n > nf ? (1.0 - p) / (double)(n - nf) : 0.0;
		for (uint64_t s = 0; s < n; ++s)
			x0_out[s] = ((uint32_t)s & fmask) == fval ? pin : pout;
	});
}

// ---------------------------------------------------------------- table / graph

esl_status esl_table_build(esl_ctx* ctx, const esl_model* m, const double* rates, esl_graph** out)
{
	return guarded(ctx, [&] {
		require(ctx && m && rates && out, ESL_E_ARG, "null argument");
		set_device(ctx);
		auto g = std::make_unique<esl_graph>();
		const int N = (int)m->m.nodes.size();
		g->g.rates.assign(rates, rates + 2 * N);
		g->g.up_ok = g->g.down_ok = 0;
		for (int i = 0; i < N; ++i)
		{
			require(rates[2 * i] >= 0 && rates[2 * i + 1] >= 0, ESL_E_ARG, "transition rate must be non-negative");
			if (rates[2 * i] != 0.0)
				g->g.up_ok |= 1u << i;
			if (rates[2 * i + 1] != 0.0)
				g->g.down_ok |= 1u << i;
		}
		esl::table_build(ctx->c, g->g, m->m);
		*out = g.release();
	});
}

void esl_graph_destroy(esl_graph* g) { delete g; }

esl_status esl_graph_reorganize(esl_ctx* ctx, esl_graph* g)
{
	return guarded(ctx, [&] {
		require(ctx && g, ESL_E_ARG, "null argument");
		set_device(ctx);
		esl::graph_reorganize(ctx->c, g->g);
	});
}

esl_status esl_graph_import(esl_ctx* ctx, esl_graph* g, const int32_t* reordered_vertices, const int32_t* sccs_offsets,
							int64_t n_sccs, int32_t terminals_count)
{
	return guarded(ctx, [&] {
		require(ctx && g && reordered_vertices && sccs_offsets, ESL_E_ARG, "null argument");
		set_device(ctx);
		esl::graph_import(ctx->c, g->g, reordered_vertices, sccs_offsets, n_sccs, terminals_count);
	});
}

esl_status esl_graph_stage_counts(const esl_graph* g, int64_t counts[3])
{
	return guarded(nullptr, [&] {
		require(g && counts, ESL_E_ARG, "null argument");
		counts[0] = g->g.mask_passes;
		counts[1] = g->g.scc_rounds;
		counts[2] = g->g.host_syncs;
	});
}

esl_status esl_graph_query(const esl_graph* g, int64_t sizes[7])
{
	return guarded(nullptr, [&] {
		require(g && sizes, ESL_E_ARG, "null argument");
		sizes[0] = g->g.n;
		sizes[1] = g->g.nnz;
		sizes[2] = g->g.reorganized ? g->g.S : 0;
		sizes[3] = g->g.T;
		sizes[4] = g->g.n_t;
		sizes[5] = g->g.N;
		sizes[6] = g->g.H;
	});
}

esl_status esl_table_export(esl_ctx* ctx, const esl_graph* g, int32_t* indptr, int32_t* rows, int32_t* cols)
{
	return guarded(ctx, [&] {
		require(ctx && g, ESL_E_ARG, "null argument");
		set_device(ctx);
		auto& c = ctx->c;
		if (indptr)
			ESL_HIP(hipMemcpyAsync(indptr, g->g.indptr.get(), (g->g.n + 1) * sizeof(int32_t), hipMemcpyDeviceToHost,
								   c.stream));
		if (rows && g->g.nnz)
			ESL_HIP(hipMemcpyAsync(rows, g->g.rows.get(), g->g.nnz * sizeof(int32_t), hipMemcpyDeviceToHost,
								   c.stream));
		if (cols && g->g.nnz)
			ESL_HIP(hipMemcpyAsync(cols, g->g.cols.get(), g->g.nnz * sizeof(int32_t), hipMemcpyDeviceToHost,
								   c.stream));
		ESL_HIP(hipStreamSynchronize(c.stream));
	});
}

esl_status esl_graph_export(esl_ctx* ctx, const esl_graph* g, int32_t* reordered, int32_t* offsets,
							int32_t* terminals_count, int32_t* scc_ids)
{
	return guarded(ctx, [&] {
		require(ctx && g, ESL_E_ARG, "null argument");
		require(g->g.reorganized, ESL_E_STATE, "esl_graph_reorganize has not run on this graph");
		set_device(ctx);
		auto& c = ctx->c;
		if (reordered)
			ESL_HIP(hipMemcpyAsync(reordered, g->g.reordered.get(), g->g.n * sizeof(int32_t), hipMemcpyDeviceToHost,
								   c.stream));
		if (offsets)
			std::copy(g->g.h_offsets.begin(), g->g.h_offsets.end(), offsets);
		if (terminals_count)
			*terminals_count = (int32_t)g->g.T;
		if (scc_ids)
			ESL_HIP(hipMemcpyAsync(scc_ids, g->g.scc.get(), g->g.n * sizeof(int32_t), hipMemcpyDeviceToHost,
								   c.stream));
		ESL_HIP(hipStreamSynchronize(c.stream));
	});
}

esl_status esl_graph_stage_ms(const esl_graph* g, double ms[2])
{
	return guarded(nullptr, [&] {
		require(g && ms, ESL_E_ARG, "null argument");
		ms[0] = g->g.stage_ms[0];
		ms[1] = g->g.stage_ms[1];
	});
}

// ---------------------------------------------------------------- solve

void esl_solve_opts_default(esl_solve_opts* o)
{
	if (!o)
		return;
	o->precision = 0;
	o->tol = 1e-13;
	o->max_iters = 20000;
	o->restart = 0;
	o->full = 0;
	o->preconditioner = 0;
	o->graph = 0;
	o->warm_pi = nullptr;
	o->inner_rtol = 1e-6;
	o->exact_fallback = 1;
}

esl_status esl_solve_batch_device(esl_ctx* ctx, const esl_graph* g, int n_sets, const double* d_rates,
								  const double* d_x0, double* d_final, const esl_solve_opts* opts,
								  esl_solve_report* report)
{
	return guarded(ctx, [&] {
		require(ctx && g && d_rates && d_x0 && d_final, ESL_E_ARG, "null argument");
		require(n_sets >= 1, ESL_E_ARG, "n_sets must be >= 1");
		set_device(ctx);
		esl_solve_opts o;
		esl_solve_opts_default(&o);
		if (opts)
			o = *opts;
		require(o.precision == 0 || o.precision == 1, ESL_E_ARG, "precision must be 0 or 1");
		require(!o.full || n_sets == 1, ESL_E_ARG, "opts.full needs a single rate set");
		esl_solve_report rep{};
		esl::solve_args_t a{ d_rates, d_x0, d_final, n_sets, o };
		esl::solve_run(ctx->c, const_cast<esl::graph_t&>(g->g), a, rep);
		if (o.full)
			esl::solution_full(ctx->c, const_cast<esl::graph_t&>(g->g), d_rates, o);
		if (report)
			*report = rep;
	});
}

esl_status esl_solve_device(esl_ctx* ctx, const esl_graph* g, const double* d_rates, const double* d_x0,
							double* d_final_state, const esl_solve_opts* opts, esl_solve_report* report)
{
	return esl_solve_batch_device(ctx, g, 1, d_rates, d_x0, d_final_state, opts, report);
}

esl_status esl_solve(esl_ctx* ctx, const esl_graph* g, const double* rates, const double* x0, double* final_state,
					 const esl_solve_opts* opts, esl_solve_report* report)
{
	return guarded(ctx, [&] {
		require(ctx && g && x0 && final_state, ESL_E_ARG, "null argument");
		set_device(ctx);
		const auto& gg = g->g;
		const double* r = rates ? rates : gg.rates.data();
		check_rates(gg, r);
		auto& c = ctx->c;
		esl::dbuf<double> d_rates, d_x0, d_final;
		d_rates.ensure(2 * gg.N);
		d_x0.ensure(gg.n);
		d_final.ensure(gg.n);
		ESL_HIP(hipMemcpyAsync(d_rates.get(), r, 2 * gg.N * sizeof(double), hipMemcpyHostToDevice, c.stream));
		ESL_HIP(hipMemcpyAsync(d_x0.get(), x0, gg.n * sizeof(double), hipMemcpyHostToDevice, c.stream));
		esl_solve_opts o;
		esl_solve_opts_default(&o);
		if (opts)
			o = *opts;
		require(o.precision == 0 || o.precision == 1, ESL_E_ARG, "precision must be 0 or 1");
		esl::dbuf<double> d_warm;
		if (o.warm_pi)
		{
			require(gg.reorganized, ESL_E_STATE, "esl_graph_reorganize must run before a solve");
			d_warm.ensure(std::max<int64_t>(gg.n_t, 1));
			ESL_HIP(hipMemcpyAsync(d_warm.get(), o.warm_pi, gg.n_t * sizeof(double), hipMemcpyHostToDevice,
								   c.stream));
			o.warm_pi = d_warm.get();
		}
		esl_solve_report rep{};
		esl::solve_args_t a{ d_rates.get(), d_x0.get(), d_final.get(), 1, o };
		esl::solve_run(c, const_cast<esl::graph_t&>(gg), a, rep);
		if (o.full)
			esl::solution_full(c, const_cast<esl::graph_t&>(gg), d_rates.get(), o);
		ESL_HIP(hipMemcpyAsync(final_state, d_final.get(), gg.n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
		ESL_HIP(hipStreamSynchronize(c.stream));
		if (report)
			*report = rep;
	});
}

esl_status esl_solution_export(esl_ctx* ctx, const esl_graph* g, int32_t* term_indptr, int32_t* term_indices,
							   double* term_data, int64_t* nonterm_nnz, int32_t* nonterm_indptr,
							   int32_t* nonterm_indices, double* nonterm_data)
{
	return guarded(ctx, [&] {
		require(ctx && g, ESL_E_ARG, "null argument");
		const auto& s = g->g.sol;
		require(s.valid, ESL_E_STATE, "no solve with opts.full = 1 has run on this graph");
		if (term_indptr)
			std::copy(s.term_indptr.begin(), s.term_indptr.end(), term_indptr);
		if (term_indices)
			std::copy(s.term_indices.begin(), s.term_indices.end(), term_indices);
		if (term_data)
			std::copy(s.term_data.begin(), s.term_data.end(), term_data);
		if (nonterm_nnz)
			*nonterm_nnz = (int64_t)s.nonterm_indices.size();
		if (nonterm_indptr)
			std::copy(s.nonterm_indptr.begin(), s.nonterm_indptr.end(), nonterm_indptr);
		if (nonterm_indices)
			std::copy(s.nonterm_indices.begin(), s.nonterm_indices.end(), nonterm_indices);
		if (nonterm_data)
			std::copy(s.nonterm_data.begin(), s.nonterm_data.end(), nonterm_data);
	});
}

// ---------------------------------------------------------------- profiling

esl_status esl_profile(esl_ctx* ctx, int enable)
{
	return guarded(ctx, [&] {
		require(ctx, ESL_E_ARG, "null argument");
		set_device(ctx);
		ESL_HIP(hipStreamSynchronize(ctx->c.stream));
		ctx->c.prof.flush();
		if (enable)
			ctx->c.prof.stats.clear();
		ctx->c.prof.on = enable != 0;
	});
}

esl_status esl_profile_stat(esl_ctx* ctx, int index, char* name_out, int name_cap, int64_t* launches,
							double* total_ms, double* total_bytes)
{
	return guarded(ctx, [&] {
		require(ctx, ESL_E_ARG, "null argument");
		ESL_HIP(hipStreamSynchronize(ctx->c.stream));
		ctx->c.prof.flush();
		const auto& st = ctx->c.prof.stats;
		require(index >= 0 && index < (int)st.size(), ESL_E_ARG, "no such kernel statistic");
		const auto& e = st[index];
		if (name_out && name_cap > 0)
		{
			std::strncpy(name_out, e.first.c_str(), name_cap - 1);
			name_out[name_cap - 1] = 0;
		}
		if (launches)
			*launches = e.second.launches;
		if (total_ms)
			*total_ms = e.second.ms;
		if (total_bytes)
			*total_bytes = e.second.bytes;
	});
}

// ---------------------------------------------------------------- plans

esl_status esl_plan_create(esl_ctx* ctx, const esl_graph* g, const double* d_x0, esl_plan** out)
{
	return guarded(ctx, [&] {
		require(ctx && g && d_x0 && out, ESL_E_ARG, "null argument");
		set_device(ctx);
		auto pl = std::make_unique<esl_plan>();
		pl->ctx = ctx;
		pl->graph = g;
		pl->p = esl::plan_create(ctx->c, const_cast<esl::graph_t&>(g->g), d_x0);
		*out = pl.release();
	});
}

void esl_plan_destroy(esl_plan* p) { delete p; }

esl_status esl_plan_query(const esl_plan* p, int64_t info[8])
{
	return guarded(nullptr, [&] {
		require(p && info, ESL_E_ARG, "null argument");
		esl::plan_info(p->p, info);
	});
}

esl_status esl_plan_solve(esl_ctx* ctx, esl_plan* p, int n_sets, const double* d_rates, double* d_final,
						  const esl_solve_opts* opts, esl_solve_report* report)
{
	return guarded(ctx, [&] {
		require(ctx && p && d_rates && d_final, ESL_E_ARG, "null argument");
		require(p->ctx == ctx, ESL_E_ARG, "plan belongs to another context");
		require(n_sets >= 1, ESL_E_ARG, "n_sets must be >= 1");
		set_device(ctx);
		esl_solve_opts o;
		esl_solve_opts_default(&o);
		if (opts)
			o = *opts;
		require(o.precision == 0 || o.precision == 1, ESL_E_ARG, "precision must be 0 or 1");
		require(!o.full, ESL_E_ARG, "opts.full is not available on plans; use esl_solve");
		esl_solve_report rep{};
		esl::plan_solve(ctx->c, p->p, n_sets, d_rates, d_final, o, rep);
		if (report)
			*report = rep;
	});
}

esl_status esl_plan_solve_noalloc(esl_ctx* ctx, esl_plan* p, int n_sets, const double* d_rates, double* d_final,
								  const esl_solve_opts* opts, esl_solve_report* report)
{
	return guarded(ctx, [&] {
		require(ctx && p && d_rates && d_final, ESL_E_ARG, "null argument");
		require(p->ctx == ctx, ESL_E_ARG, "plan belongs to another context");
		require(n_sets >= 1, ESL_E_ARG, "n_sets must be >= 1");
		set_device(ctx);
		esl_solve_opts o;
		esl_solve_opts_default(&o);
		if (opts)
			o = *opts;
		require(o.precision == 0 || o.precision == 1, ESL_E_ARG, "precision must be 0 or 1");
		require(!o.full, ESL_E_ARG, "opts.full is not available on plans; use esl_solve");
		o.graph = 0; // never captured by the library
		esl_solve_report rep{};
		esl::plan_solve_noalloc(ctx->c, p->p, n_sets, d_rates, d_final, o, rep);
		if (report)
			*report = rep;
	});
}

esl_status esl_plan_gather_terminal(esl_ctx* ctx, const esl_plan* p, int n_sets, const double* d_final,
									double* d_out)
{
	return guarded(ctx, [&] {
		require(ctx && p && d_final && d_out, ESL_E_ARG, "null argument");
		require(p->ctx == ctx, ESL_E_ARG, "plan belongs to another context");
		require(n_sets >= 1, ESL_E_ARG, "n_sets must be >= 1");
		set_device(ctx);
		esl::plan_gather_terminal(ctx->c, p->p, n_sets, d_final, d_out);
	});
}

esl_status esl_plan_set_results(const esl_plan* p, int n_sets, double* residual, int32_t* iterations,
								int32_t* method)
{
	return guarded(nullptr, [&] {
		require(p != nullptr, ESL_E_ARG, "null plan");
		require(n_sets >= 1, ESL_E_ARG, "n_sets must be >= 1");
		esl::plan_set_results(p->p, n_sets, residual, iterations, method);
	});
}

esl_status esl_plan_set_times(esl_ctx* ctx, const esl_plan* p, int n_sets, double* start_ms, double* end_ms)
{
	return guarded(ctx, [&] {
		require(ctx && p, ESL_E_ARG, "null argument");
		require(n_sets >= 1, ESL_E_ARG, "n_sets must be >= 1");
		esl::plan_set_times(p->p, n_sets, ctx->c.wall_khz, start_ms, end_ms);
	});
}

esl_status esl_plan_set_schedule(esl_plan* p, int schedule, int restart)
{
	return guarded(nullptr, [&] {
		require(p != nullptr, ESL_E_ARG, "null plan");
		require(schedule >= -1 && schedule <= 5, ESL_E_ARG, "schedule must be -1 or 0..5");
		esl::plan_set_schedule(p->p, schedule, restart);
	});
}

esl_status esl_plan_autotune(esl_ctx* ctx, esl_plan* p, int n_sets, const double* d_rates, double* d_final,
							 const esl_solve_opts* opts, int32_t choice[2], double* best_ms)
{
	return guarded(ctx, [&] {
		require(ctx && p && d_rates && d_final && choice && best_ms, ESL_E_ARG, "null argument");
		require(p->ctx == ctx, ESL_E_ARG, "plan belongs to another context");
		require(n_sets >= 1, ESL_E_ARG, "n_sets must be >= 1");
		set_device(ctx);
		esl_solve_opts o;
		esl_solve_opts_default(&o);
		if (opts)
			o = *opts;
		require(o.precision == 0 || o.precision == 1, ESL_E_ARG, "precision must be 0 or 1");
		require(!o.full, ESL_E_ARG, "opts.full is not available on plans; use esl_solve");
		esl::plan_autotune(ctx->c, p->p, n_sets, d_rates, d_final, o, choice, best_ms);
	});
}

esl_status esl_plan_solve_many(esl_ctx* ctx, int n_plans, esl_plan* const* plans, const int* n_sets,
							   const double* const* d_rates, double* const* d_final, const esl_solve_opts* opts,
							   esl_solve_report* reports)
{
	return guarded(ctx, [&] {
		require(ctx && plans && n_sets && d_rates && d_final && reports, ESL_E_ARG, "null argument");
		require(n_plans >= 1, ESL_E_ARG, "n_plans must be >= 1");
		std::vector<void*> ps(n_plans);
		for (int i = 0; i < n_plans; ++i)
		{
			require(plans[i] && d_rates[i] && d_final[i], ESL_E_ARG, "null plan or buffer");
			require(plans[i]->ctx == ctx, ESL_E_ARG, "plan belongs to another context");
			require(n_sets[i] >= 1, ESL_E_ARG, "n_sets must be >= 1");
			for (int j = 0; j < i; ++j)
				require(plans[j] != plans[i], ESL_E_ARG, "a plan appears twice (its buffers would be shared)");
			ps[i] = plans[i]->p;
		}
		set_device(ctx);
		esl_solve_opts o;
		esl_solve_opts_default(&o);
		if (opts)
			o = *opts;
		require(o.precision == 0 || o.precision == 1, ESL_E_ARG, "precision must be 0 or 1");
		require(!o.full, ESL_E_ARG, "opts.full is not available on plans; use esl_solve");
		require(o.warm_pi == nullptr, ESL_E_ARG, "warm starts are per plan: use esl_plan_solve");
		esl::plan_solve_many(ctx->c, n_plans, ps.data(), n_sets, d_rates, d_final, o, reports);
	});
}

// ---------------------------------------------------------------- memory

esl_status esl_device_alloc(esl_ctx* ctx, size_t bytes, void** out)
{
	return guarded(ctx, [&] {
		require(ctx && out, ESL_E_ARG, "null argument");
		set_device(ctx);
		ESL_HIP(hipMalloc(out, bytes ? bytes : 1));
	});
}

esl_status esl_device_free(esl_ctx* ctx, void* p)
{
	return guarded(ctx, [&] {
		require(ctx, ESL_E_ARG, "null argument");
		set_device(ctx);
		if (p)
			ESL_HIP(hipFree(p));
	});
}

esl_status esl_memcpy_h2d(esl_ctx* ctx, void* dst, const void* src, size_t bytes)
{
	return guarded(ctx, [&] {
		require(ctx && dst && src, ESL_E_ARG, "null argument");
		set_device(ctx);
		ESL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->c.stream));
		ESL_HIP(hipStreamSynchronize(ctx->c.stream));
	});
}

esl_status esl_memcpy_d2h(esl_ctx* ctx, void* dst, const void* src, size_t bytes)
{
	return guarded(ctx, [&] {
		require(ctx && dst && src, ESL_E_ARG, "null argument");
		set_device(ctx);
		ESL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->c.stream));
		ESL_HIP(hipStreamSynchronize(ctx->c.stream));
	});
}

esl_status esl_memcpy_d2d(esl_ctx* ctx, void* dst, const void* src, size_t bytes)
{
	return guarded(ctx, [&] {
		require(ctx && dst && src, ESL_E_ARG, "null argument");
		set_device(ctx);
		esl::device_copy(ctx->c, dst, src, bytes);
		ESL_HIP(hipStreamSynchronize(ctx->c.stream));
	});
}

} // extern "C"
