// Internal declarations shared by the host API and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/esl.h"
#include "model.hpp"

#define ESL_HIP(call)                                                                                          \
	do                                                                                                         \
	{                                                                                                          \
		hipError_t e_ = (call);                                                                                \
		if (e_ != hipSuccess)                                                                                  \
			throw ::esl::error(e_ == hipErrorOutOfMemory ? ESL_E_OOM : ESL_E_DEVICE,                           \
							   std::string(#call) + ": " + hipGetErrorString(e_) + " (" + __FILE__ + ":" +    \
								   std::to_string(__LINE__) + ")");                                            \
	} while (0)

namespace esl {

// Owning device buffer (hipMallocAsync-free on purpose: plain hipMalloc/hipFree).
template <typename T>
struct dbuf
{
	T* p = nullptr;
	size_t n = 0;
	dbuf() = default;
	dbuf(const dbuf&) = delete;
	dbuf& operator=(const dbuf&) = delete;
	dbuf(dbuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr, o.n = 0; }
	dbuf& operator=(dbuf&& o) noexcept
	{
		reset();
		p = o.p, n = o.n;
		o.p = nullptr, o.n = 0;
		return *this;
	}
	~dbuf() { reset(); }
	void reset()
	{
		if (p)
			(void)hipFree(p);
		p = nullptr, n = 0;
	}
	// grows only; contents are not preserved
	void ensure(size_t count)
	{
		if (count <= n && p)
			return;
		reset();
		if (count == 0)
			count = 1;
		ESL_HIP(hipMalloc(&p, count * sizeof(T)));
		n = count;
	}
	T* get() const { return p; }
};

// Device form of the rules: per node {k, input offset, table bit offset}.
struct rule_dev_t
{
	int32_t k;
	int32_t in_off;
	int32_t tab_off; // in bits
	int32_t pad;
};

struct graph_t;

// Per-kernel timing with HIP events on the context stream (off by default).
struct kstat_t
{
	int64_t launches = 0;
	double ms = 0.0;
	double bytes = 0.0; // algorithmic bytes summed over the launches
};

struct prof_t
{
	bool on = false;
	struct pending_t
	{
		const char* name;
		hipEvent_t a, b;
		double bytes;
	};
	std::vector<pending_t> pending;
	std::vector<hipEvent_t> pool;
	size_t used = 0;
	std::vector<std::pair<std::string, kstat_t>> stats;
	// while a HIP graph is being captured: its event pairs go to the graph's owner
	std::vector<hipEvent_t>* capture_events = nullptr;
	std::vector<pending_t>* capture_pending = nullptr;

	hipEvent_t next_event()
	{
		if (capture_events)
		{
			hipEvent_t e;
			if (hipEventCreate(&e) != hipSuccess)
				throw error(ESL_E_DEVICE, "hipEventCreate failed");
			capture_events->push_back(e);
			return e;
		}
		if (used == pool.size())
		{
			hipEvent_t e;
			if (hipEventCreate(&e) != hipSuccess)
				throw error(ESL_E_DEVICE, "hipEventCreate failed");
			pool.push_back(e);
		}
		return pool[used++];
	}
	kstat_t& stat(const std::string& name)
	{
		for (auto& p : stats)
			if (p.first == name)
				return p.second;
		stats.emplace_back(name, kstat_t{});
		return stats.back().second;
	}
	// books the event pairs of one completed launch sequence (the stream must be idle)
	void book(const std::vector<pending_t>& list)
	{
		for (auto& p : list)
		{
			float ms = 0;
			if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess)
			{
				kstat_t& k = stat(p.name);
				k.launches++;
				k.ms += ms;
				k.bytes += p.bytes;
			}
			else
				(void)hipGetLastError(); // an untimed pair is skipped, not left as the next call's error
		}
	}
	void flush()
	{
		book(pending);
		pending.clear();
		used = 0;
	}
	~prof_t()
	{
		for (auto e : pool)
			(void)hipEventDestroy(e);
	}
};

struct ctx_t
{
	int device = 0;
	int lds_per_block = 65536; // dynamic + static LDS a workgroup may use on this device
	hipStream_t stream = nullptr;
	std::string last_error;
	hipEvent_t ev[16] = {};
	// reusable scratch
	dbuf<uint8_t> scratch_a, scratch_b;
	dbuf<int32_t> flags;	// small device counters / flags
	int32_t* h_flags = nullptr; // pinned mirror
	// pinned staging of the per-set outputs of a launch, read back in one copy
	// (allocated with the context: calls that may not allocate use it too)
	unsigned char* h_stage = nullptr;
	size_t h_stage_bytes = 0;
	prof_t prof;
	// forked streams (default priority) and fork/join events for plans solved concurrently
	std::vector<hipStream_t> aux;
	std::vector<hipEvent_t> aux_ev;
	// stream (default priority) and fork/join events of a plan's final-vector fill
	hipStream_t fill_s = nullptr;
	hipEvent_t fill_ev[2] = {};
	// captured sequence of the last concurrent multi-plan solve (owned by solve.hip)
	void* multi = nullptr;
	void (*multi_free)(void*) = nullptr;
	// device address of the library's static scratch array on this device (noalloc_init)
	void* noalloc_scratch = nullptr;
	double wall_khz = 100000.0; // rate of the device wall clock (wall_clock64) in kHz
};

// size of the static device scratch behind esl_plan_solve_noalloc, and the
// device indices it keeps an address for
constexpr size_t noalloc_scratch_bytes = (size_t)64 << 20;
constexpr int max_devices = 64;

// Appends an event-record node to the graph being captured on `s`, after the
// nodes captured so far, and makes it the dependency of what follows.
inline void record_event_node(hipStream_t s, hipEvent_t e)
{
	hipStreamCaptureStatus st;
	unsigned long long id = 0;
	hipGraph_t g = nullptr;
	const hipGraphNode_t* deps = nullptr;
	size_t nd = 0;
	ESL_HIP(hipStreamGetCaptureInfo_v2(s, &st, &id, &g, &deps, &nd));
	if (st != hipStreamCaptureStatusActive)
		throw error(ESL_E_STATE, "stream is not capturing");
	hipGraphNode_t node;
	ESL_HIP(hipGraphAddEventRecordNode(&node, g, deps, nd, e));
	ESL_HIP(hipStreamUpdateCaptureDependencies(s, &node, 1, hipStreamSetCaptureDependencies));
}

// Launch `launch` (a lambda issuing one kernel on c.stream); with profiling on,
// bracket it with events and book `bytes` algorithmic bytes under `name`.
template <typename F>
inline void timed_launch(ctx_t& c, const char* name, double bytes, F&& launch)
{
	if (!c.prof.on)
	{
		launch();
		return;
	}
	hipEvent_t a = c.prof.next_event(), b = c.prof.next_event();
	if (c.prof.capture_pending)
	{
		// an hipEventRecord captured from the stream cannot be timed on replay;
		// explicit event-record nodes can
		record_event_node(c.stream, a);
		launch();
		record_event_node(c.stream, b);
		c.prof.capture_pending->push_back({ name, a, b, bytes });
		return;
	}
	ESL_HIP(hipEventRecord(a, c.stream));
	launch();
	ESL_HIP(hipEventRecord(b, c.stream));
	c.prof.pending.push_back({ name, a, b, bytes });
	if (c.prof.pending.size() > 4096)
	{
		ESL_HIP(hipStreamSynchronize(c.stream));
		c.prof.flush();
	}
}

struct solution_t
{
	bool valid = false;
	std::vector<int32_t> term_indptr, term_indices;
	std::vector<double> term_data;
	std::vector<int32_t> nonterm_indptr, nonterm_indices;
	std::vector<double> nonterm_data;
};

struct graph_t
{
	int N = 0;
	int64_t n = 0, nnz = 0;
	std::vector<double> rates; // rates given to esl_table_build (zero pattern)
	uint32_t up_ok = 0, down_ok = 0;

	// table
	dbuf<rule_dev_t> rules;
	dbuf<int32_t> rule_inputs;
	dbuf<uint32_t> rule_tables;
	dbuf<uint32_t> mask;   // n: bit i set iff s -> s ^ (1 << i)
	dbuf<int32_t> indptr;  // n + 1
	dbuf<int32_t> rows;	   // nnz (targets)
	dbuf<int32_t> cols;	   // nnz (sources)

	// graph
	bool reorganized = false;
	int64_t S = 0, T = 0, n_t = 0;
	int32_t H = 0;
	dbuf<int32_t> scc;		 // n: canonical SCC id (minimum vertex)
	dbuf<int32_t> height;	 // n: height of the vertex's SCC
	dbuf<int32_t> reordered; // n
	dbuf<int32_t> offsets;	 // S + 1
	dbuf<int32_t> rank;		 // n: position of the vertex's SCC in the canonical SCC order
	std::vector<int32_t> h_offsets; // host copy of offsets (S + 1)

	double stage_ms[2] = { 0, 0 };
	// graph stage counters: passes over the flip masks, SCC rounds, host round trips
	int64_t mask_passes = 0, scc_rounds = 0, host_syncs = 0;
	solution_t sol;
};

// ---- kernel launchers (defined in the .hip files) ----
void table_build(ctx_t& c, graph_t& g, const model_t& m);
void device_copy(ctx_t& c, void* dst, const void* src, size_t bytes);
void graph_reorganize(ctx_t& c, graph_t& g);
void graph_import(ctx_t& c, graph_t& g, const int32_t* h_reordered, const int32_t* h_offsets, int64_t S, int64_t T);

struct solve_args_t
{
	const double* d_rates; // [n_sets][2N]
	const double* d_x0;	   // [n] shared
	double* d_final;	   // [n_sets][n]
	int n_sets;
	esl_solve_opts opts;
};
void solve_run(ctx_t& c, graph_t& g, const solve_args_t& a, esl_solve_report& rep);
// rate-independent solve plans (solve.hip)
void* plan_create(ctx_t& c, graph_t& g, const double* d_x0);
void plan_destroy(void* p);
void plan_solve(ctx_t& c, void* p, int B, const double* d_rates, double* d_final, const esl_solve_opts& o,
				esl_solve_report& rep);
void plan_info(const void* p, int64_t info[8]);
void plan_solve_many(ctx_t& c, int np, void* const* plans, const int* B, const double* const* d_rates,
					 double* const* d_final, const esl_solve_opts& o, esl_solve_report* reps);
void plan_autotune(ctx_t& c, void* p, int B, const double* d_rates, double* d_final, const esl_solve_opts& o,
				   int32_t choice[2], double* best_ms);
void plan_gather_terminal(ctx_t& c, const void* p, int B, const double* d_final, double* d_out);
void plan_set_schedule(void* p, int path, int restart);
void plan_set_results(const void* p, int B, double* res, int32_t* its, int32_t* method);
void plan_set_times(const void* p, int B, double wall_khz, double* start_ms, double* end_ms);
// pinned staging buffer of a context for per-set outputs (32 B per set with the
// clocks: batches of up to 8192 sets in one copy)
constexpr size_t stage_bytes = 256 * 1024;
// the plan solve whose buffers are slices of the static scratch (no allocation)
void noalloc_init(ctx_t& c); // once per device, with that device current
void plan_solve_noalloc(ctx_t& c, void* p, int B, const double* d_rates, double* d_final, const esl_solve_opts& o,
						esl_solve_report& rep);
// the implementation object behind an esl_ctx handle (defined in esl_api.cpp)
ctx_t& ctx_impl(esl_ctx* ctx);
// solution_term / solution_nonterm of one rate set into g.sol (reference layout)
void solution_full(ctx_t& c, graph_t& g, const double* d_rates, const esl_solve_opts& opts);

// device scan helper: out[i] = sum_{j<i} in[j] (exclusive), out[n] = total; in/out int32
void exclusive_scan_i32(ctx_t& c, const int32_t* in, int32_t* out, int64_t n);

} // namespace esl
