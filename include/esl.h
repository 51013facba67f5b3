/*
 * esl.h -- C ABI of the MI355X (gfx950) stationary-distribution pipeline for
 * asynchronous Boolean networks.
 *
 * One shared library (libesl.so), extern "C", plain pointers and sizes.  No
 * C++ exception and no exit() crosses this boundary: every entry point returns
 * an esl_status and leaves a message readable through esl_last_error().
 *
 * The entry points replace the C++ class API of the reference
 * (asmelko/paraExaStoLog, the headers in cuda/src); each declaration names the reference
 * interface it stands in for.  Terminology follows the reference: a model has
 * N nodes, a state is an N-bit integer (bit i = node i in file order), the
 * transition table is a CSC matrix over source states (indptr / rows = target
 * state / cols = source state), the graph stage yields reordered_vertices,
 * sccs_offsets and terminals_count, and the solver yields the dense
 * final_state plus solution_term (CSC) and solution_nonterm (CSR).
 *
 * Threading: one esl_ctx owns one device and one HIP stream.  Calls on one ctx
 * must be serialised by the caller; separate ctxs may run concurrently.  Every
 * call is synchronous on return unless its name ends in _async.
 */
#ifndef ESL_H
#define ESL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct esl_ctx esl_ctx;
typedef struct esl_model esl_model;
typedef struct esl_graph esl_graph;

typedef enum
{
	ESL_OK = 0,
	ESL_E_IO = 1,		/* file missing / unreadable */
	ESL_E_PARSE = 2,	/* malformed .bnet line or formula */
	ESL_E_ARG = 3,		/* bad argument: unknown node name, negative rate, size mismatch */
	ESL_E_DEVICE = 4,	/* HIP runtime error, no device, or the library was built without one */
	ESL_E_NUMERIC = 5,	/* a solve did not reach its tolerance (after the exact fallback) */
	ESL_E_OOM = 6,		/* device or host allocation failed */
	ESL_E_STATE = 7		/* call order violated (e.g. export before build) */
} esl_status;

/* ---- context ------------------------------------------------------------------
 * Replaces cu_context (cuda/src/cu_context.h:15-23, cu_context.cpp:34-46):
 * owns the device and the stream instead of cuSPARSE/cuSOLVER handles.      */
esl_status esl_ctx_create(int device, esl_ctx** out);
void esl_ctx_destroy(esl_ctx* ctx);
/* Message of the last failed call on ctx; ctx == NULL gives the calling
 * thread's last message of the context-free calls (model, rates, x0). */
const char* esl_last_error(const esl_ctx* ctx);
const char* esl_version(void);

/* ---- model --------------------------------------------------------------------
 * Replaces model_builder::construct_model (cuda/src/model.h:41,
 * model.cpp:80-125) with the formula grammar of BoolExprParser
 * (cuda/src/boolstuff/BoolExprParser.cpp:64-200).  Each rule is kept as a
 * truth table over the nodes it reads.  Models of more than 26 nodes (2^N
 * states as 32-bit integers, the table in int32) are ESL_E_ARG, rules reading
 * more than 24 nodes ESL_E_PARSE. */
esl_status esl_model_load(const char* bnet_path, esl_model** out);
esl_status esl_model_parse(const char* bnet_text, esl_model** out);
void esl_model_destroy(esl_model* m);
int esl_model_num_nodes(const esl_model* m);
const char* esl_model_node_name(const esl_model* m, int i);
/* inputs of node i (ascending node indices) and its truth table (bit k of the
 * table index = k-th input); *k_out = number of inputs.  table_bits holds
 * 2^k bytes (0/1).  Either output pointer may be NULL to query k only. */
esl_status esl_model_rule(const esl_model* m, int i, int* k_out, int32_t* inputs, uint8_t* table_bits);

/* ---- transition rates -------------------------------------------------------
 * Replaces transition_rates::{generate_uniform, generate_normal}
 * (cuda/src/transition_rates.h:26-30, transition_rates.cpp:7-97).
 * Layout rates[2i] = up rate of node i, rates[2i+1] = down rate.
 * kind 0 = UNIFORM (every rate 1.0), kind 1 = NORMAL (std::default_random_engine
 * + std::normal_distribution<float>(mean, std), negative draws redrawn up to
 * 1000 times, generated node by node, up before down, only for rates that are
 * not overridden).  Overrides are matched by node name; an unknown name or a
 * negative value is ESL_E_ARG. */
esl_status esl_rates_fill(const esl_model* m, int kind, double mean, double std_dev, const char* const* up_names,
						  const double* up_vals, int n_up, const char* const* down_names, const double* down_vals,
						  int n_down, double* rates_out /* 2N */);

/* ---- initial state ----------------------------------------------------------
 * Replaces initial_state::initial_state (cuda/src/initial_state.h:9-10,
 * initial_state.cu:7-60): x0[s] = p/|F| on the |F| states matching every fixed
 * (node, value) pair and (1-p)/(2^N-|F|) elsewhere; n_fixed == 0 gives the
 * uniform 1/2^N.  x0_out holds 2^N doubles. */
esl_status esl_initial_state(const esl_model* m, const char* const* fixed_names, const uint8_t* fixed_vals,
							 int n_fixed, double p, double* x0_out);

/* ---- transition table -------------------------------------------------------
 * Replaces transition_table(ctx, model, rates).construct_table()
 * (cuda/src/transition_table.h:21,25; transition_table.cu:121-174).  Builds on
 * the device, from the rule truth tables, the per-state flip mask and the
 * CSC table over source states.  A zero rate removes that direction.  The
 * rates only matter through their zero pattern; they are kept for esl_solve
 * when it is called with rates == NULL. */
esl_status esl_table_build(esl_ctx* ctx, const esl_model* m, const double* rates /* 2N */, esl_graph** out);
void esl_graph_destroy(esl_graph* g);

/* ---- transition graph -------------------------------------------------------
 * Replaces transition_graph(ctx, rows, cols, indptr).reorganize_vertices()
 * (cuda/src/transition_graph.h:21-27, transition_graph.cu:123-202):
 * SCC decomposition, heights of the condensation DAG and the canonical order
 * [terminal SCCs by ascending id | transient SCCs by (height descending, id
 * descending)], vertices ascending inside an SCC, SCC id = minimum vertex. */
esl_status esl_graph_reorganize(esl_ctx* ctx, esl_graph* g);

/* The graph stage from stored results instead of recomputing it: the
 * reordered_vertices (n), sccs_offsets (n_sccs + 1) and terminals_count of an
 * earlier esl_graph_export of the same table (persistent_data::ordered_vertices,
 * terminals_offsets ++ nonterminals_offsets, cuda/src/persistent_solution.h:5-24,
 * loaded by persistent_solution::deserialize, persistent_solution.cu:104-199, and
 * used by solver(ctx, persisted, rates, state), solver.cu:25-51).  Checked to be a
 * permutation with ascending blocks; the SCC structure itself is trusted, as the
 * reference trusts its file.  Heights are not restored (query sizes[6] = 0). */
esl_status esl_graph_import(esl_ctx* ctx, esl_graph* g, const int32_t* reordered_vertices, const int32_t* sccs_offsets,
							int64_t n_sccs, int32_t terminals_count);

/* sizes[0] = n (states), [1] = nnz (transitions), [2] = S (SCCs, 0 before
 * reorganize), [3] = T (terminal SCCs), [4] = n_t (states in terminal SCCs),
 * [5] = N (nodes), [6] = max SCC height. */
esl_status esl_graph_query(const esl_graph* g, int64_t sizes[7]);

/* Counters of the last esl_graph_reorganize: counts[0] = passes over the
 * per-state flip masks (trimming, colouring and closure passes and height
 * levels), [1] = SCC rounds, [2] = host round trips.  No reference
 * counterpart (diagnostics of the graph stage). */
esl_status esl_graph_stage_counts(const esl_graph* g, int64_t counts[3]);

/* Table members indptr (n+1), rows (nnz, target states), cols (nnz, source
 * states); any pointer may be NULL. */
esl_status esl_table_export(esl_ctx* ctx, const esl_graph* g, int32_t* indptr, int32_t* rows, int32_t* cols);

/* Graph members reordered_vertices (n), sccs_offsets (S+1), terminals_count;
 * also the canonical SCC id of every state (n) if scc_ids != NULL. */
esl_status esl_graph_export(esl_ctx* ctx, const esl_graph* g, int32_t* reordered_vertices, int32_t* sccs_offsets,
							int32_t* terminals_count, int32_t* scc_ids);

/* ---- solver -------------------------------------------------------------------
 * Replaces solver(ctx, table, graph, rates, state).solve() and its members
 * final_state / solution_term / solution_nonterm (cuda/src/solver.h:40-64,
 * solver.cu:102-334). */

typedef struct
{
	int precision;		 /* 0 = FP64 Krylov solves; 1 = FP32 inner solves + FP64 residual refinement */
	double tol;			 /* relative residual target of every linear solve (default 1e-13) */
	int max_iters;		 /* Krylov iterations per solve before ESL_E_NUMERIC (default 20000) */
	int restart;		 /* GMRES restart length: at most 64 for the per-set solvers, 128 for the
						  * device-wide solver; 0 (default): a plan's autotuned length, else 32 (the
						  * device-wide solver halves it while its FP64 basis would exceed half the
						  * free device memory) */
	int full;			 /* 1: also build solution_term / solution_nonterm (esl_solution_export) */
	int preconditioner;	 /* 0 = default (ILU(0)), 1 = Jacobi, 2 = ILU(0) (level-scheduled) */
	int graph;			 /* 1: capture the numeric sequence of a plan in a HIP graph and replay it */
	const double* warm_pi; /* optional warm start of the terminal kernels: predicted final-vector values
							* on the terminal states in canonical order (reordered_vertices[0, n_t)),
							* [n_sets][n_t]; host memory for esl_solve, device memory otherwise */
	double inner_rtol;	 /* precision 1: residual reduction one FP32 cycle aims for before the FP64
						  * refinement step (default 1e-6; 0 keeps iterating towards tol) */
	int exact_fallback;	 /* 1 (default): a rate set still above tol after its Krylov solves is solved
						  * exactly -- FP64 dense LU with partial pivoting of each diagonal block of its
						  * system (blocks up to 16384 unknowns) plus two refinement steps -- instead of
						  * failing the call; 0: off; 2: every set is solved that way (a check) */
} esl_solve_opts;

typedef struct
{
	double residual;	 /* max relative FP64 residual over the linear systems solved */
	int iterations;		 /* total Krylov iterations */
	int refinements;	 /* FP64 refinement steps (precision 1) */
	int64_t reachable;	 /* states reachable from supp(x0) */
	int64_t unknowns;	 /* unknowns of the linear systems (terminal kernels + ambiguous transient states) */
	int terminals_reached; /* terminal SCCs reachable from supp(x0) */
	int sets_iterated;	 /* rate sets whose initial iterate (zero, or the warm start) missed the tolerance */
	int fp64_fallback;	 /* precision 1: rate sets whose FP32 cycles stalled and that were finished in FP64 */
	double ms_total;	 /* device time of the call */
	double ms_stage[8];	 /* reach, terminal masks, setup, terminal solve, transient solve, final, -, - */
	int64_t iterations_sum; /* Krylov iterations summed over the rate sets */
	int exact_solves;	 /* rate sets solved by the exact dense LU (esl_solve_opts.exact_fallback) */
} esl_solve_report;

void esl_solve_opts_default(esl_solve_opts* o);

/* final_state (2^N doubles, host) = lim exp(K t) x0.  rates == NULL reuses the
 * rates given to esl_table_build; otherwise their zero pattern must equal it
 * (persistent_solution::has_compatible_zero_rates,
 * cuda/src/persistent_solution.cu:225-235), else ESL_E_ARG. */
esl_status esl_solve(esl_ctx* ctx, const esl_graph* g, const double* rates, const double* x0, double* final_state,
					 const esl_solve_opts* opts, esl_solve_report* report);

/* Same with device pointers (rates: 2N, x0 and final_state: 2^N doubles). */
esl_status esl_solve_device(esl_ctx* ctx, const esl_graph* g, const double* d_rates, const double* d_x0,
							double* d_final_state, const esl_solve_opts* opts, esl_solve_report* report);

/* solution_term (CSC n x T: indptr T+1, indices n_t, data n_t) and
 * solution_nonterm (CSR T x n: indptr T+1, indices nnz_L, data nnz_L) of the
 * last esl_solve on g with opts.full = 1, in the reference layout
 * (solver.cu:108-111, 291-299).  Query nnz_L with all pointers NULL. */
esl_status esl_solution_export(esl_ctx* ctx, const esl_graph* g, int32_t* term_indptr, int32_t* term_indices,
							   double* term_data, int64_t* nonterm_nnz, int32_t* nonterm_indptr,
							   int32_t* nonterm_indices, double* nonterm_data);

/* ---- batched solves -----------------------------------------------------------
 * Many rate sets of one topology (same zero pattern) share the table, the
 * graph, reachability and the system structure (the reference's counterpart is
 * the .symb refactor path: solver(ctx, persisted, rates, state),
 * cuda/src/solver.cu:25-51 and linear_system_solve.cu:327-490).
 * Device pointers: d_rates [n_sets][2N], d_x0 [2^N] shared by all sets,
 * d_final [n_sets][2^N] (set-major).  report may be NULL. */
esl_status esl_solve_batch_device(esl_ctx* ctx, const esl_graph* g, int n_sets, const double* d_rates,
								  const double* d_x0, double* d_final, const esl_solve_opts* opts,
								  esl_solve_report* report);

/* ---- plans: the rate-independent part of a solve, reused across rate sets ---
 * A plan fixes one graph and one initial state (copied) and holds everything a
 * solve needs that does not depend on the rates: reachability, which terminal
 * SCCs each state can reach, the unknowns of the linear systems and their
 * coupling structure and popcount levels.  The reference's counterpart is the
 * symbolic reuse of a .symb file (solver(ctx, persisted, rates, state),
 * cuda/src/solver.cu:25-51; refactor, linear_system_solve.cu:327-490).
 * Every rate set solved with a plan must have the table's zero pattern. */
typedef struct esl_plan esl_plan;
esl_status esl_plan_create(esl_ctx* ctx, const esl_graph* g, const double* d_x0, esl_plan** out);
void esl_plan_destroy(esl_plan* p);
/* info[0] unknowns per set, [1] terminal-kernel unknowns, [2] ambiguous transient
 * states, [3] reachable terminal SCCs, [4] terminal SCCs with >= 2 states,
 * [5] coupling nonzeros, [6] n_t, [7] reserved */
esl_status esl_plan_query(const esl_plan* p, int64_t info[8]);
/* n_sets rate sets: d_rates [n_sets][2N], d_final [n_sets][2^N] (device).  With
 * opts.graph = 1 the numeric sequence is captured in a HIP graph on the first
 * call and replayed while n_sets, the pointers and the options stay the same. */
esl_status esl_plan_solve(esl_ctx* ctx, esl_plan* p, int n_sets, const double* d_rates, double* d_final,
						  const esl_solve_opts* opts, esl_solve_report* report);

/* esl_plan_solve without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory during a
 * solve (tight serving loops, harnesses that forbid hipMalloc).  Every per-run
 * buffer (anchors, coefficients, iterates, Krylov basis, partial sums) is a
 * slice of one static device array of 64 MiB that the library holds on each
 * GPU until the process ends; the caller passes the plan (made beforehand by
 * esl_plan_create) and the input and output buffers as for esl_plan_solve.
 * A batch whose buffers exceed 64 MiB (about 200 rate sets of a 1,000-unknown
 * system at restart 32), or a plan that needs the device-wide solver, is
 * ESL_E_ARG with the needed size in the message.  opts.graph is ignored (no
 * capture) and kernel profiling is paused during the call.  Calls on one GPU
 * share the array: calls that overlap in time on one GPU (two threads, two
 * contexts) race; each call writes every slice it reads before reading it. */
esl_status esl_plan_solve_noalloc(esl_ctx* ctx, esl_plan* p, int n_sets, const double* d_rates, double* d_final,
								  const esl_solve_opts* opts, esl_solve_report* report);

/* d_out[n_sets][n_t] = d_final[set][reordered_vertices[i]] for i < n_t: the final
 * vectors on the terminal states in canonical order -- the layout of opts->warm_pi
 * and of the surrogate's targets (include/esl_mlp.h).  No reference counterpart. */
esl_status esl_plan_gather_terminal(esl_ctx* ctx, const esl_plan* p, int n_sets, const double* d_final,
									double* d_out);

/* Autotuner: runs this batch under each candidate schedule -- 0 per-set solve with
 * the system in LDS (FP32 basis in registers where it fits), 1 per-set from global
 * memory, 2 device-wide, 3 per-set with the system in LDS and the basis in global
 * memory, 4 and 5 as 0 with 3 and 2 workgroups per CU (more registers: the basis of
 * systems up to 1024 unknowns then stays in registers) -- and restart length
 * 16 / 24 / 32 / 64, keeps the fastest that meets opts->tol in the plan, and returns
 * it (choice[0] = schedule 0..5, choice[1] = restart; *best_ms = its time per
 * batch).
 * d_final receives the batch's results.  Later esl_plan_solve / esl_plan_solve_many
 * calls with opts->restart == 0 use the stored choice (esl_plan_query info[7] =
 * 1000 * schedule + restart, or -1 before tuning).  No reference counterpart. */
esl_status esl_plan_autotune(esl_ctx* ctx, esl_plan* p, int n_sets, const double* d_rates, double* d_final,
							 const esl_solve_opts* opts, int32_t choice[2], double* best_ms);

/* Per-set outcome of the last solve on p (esl_plan_solve, esl_plan_solve_many,
 * esl_plan_solve_noalloc, esl_plan_autotune), n_sets as in that call; any output
 * may be NULL.  residual[i]: relative FP64 residual of set i's systems;
 * iterations[i]: its Krylov iterations; method[i]: 0 Krylov solve at the asked
 * precision, 1 finished by FP64 Krylov after its FP32 cycles stalled, 2 exact
 * dense LU (esl_solve_opts.exact_fallback), -1 above the tolerance (the call
 * returned ESL_E_NUMERIC; the other sets' results are valid).  The reference
 * solves one rate set per call and exits on failure (cu_context.cpp:7-32), so
 * there is no counterpart there. */
esl_status esl_plan_set_results(const esl_plan* p, int n_sets, double* residual, int32_t* iterations,
								int32_t* method);

/* When each set's workgroup of the last per-set launch on p started and ended,
 * in ms after the earliest start (device wall clock): the spread of per-set
 * solve times and how long the launch waited for its slowest set.  ESL_E_STATE
 * when that solve ran device-wide or with another n_sets.  Any output may be
 * NULL.  No reference counterpart. */
esl_status esl_plan_set_times(esl_ctx* ctx, const esl_plan* p, int n_sets, double* start_ms, double* end_ms);

/* Pins the schedule esl_plan_autotune would store (e.g. one found earlier):
 * schedule 0..5 as above with the given restart length, or -1 to forget it. */
esl_status esl_plan_set_schedule(esl_plan* p, int schedule, int restart);

/* Several plans (distinct networks, e.g. of uneven size) in one call: plan i solves
 * n_sets[i] rate sets d_rates[i] -> d_final[i]; the per-set solves of all plans run
 * concurrently on forked streams (opts->graph = 1: captured once as one HIP graph and
 * replayed while plans, sizes, pointers and options stay the same).  Plans whose system
 * needs the device-wide solver run one after another first.  reports[n_plans].
 * No reference counterpart (the reference solves one network per process); it serves
 * batches over several networks. */
esl_status esl_plan_solve_many(esl_ctx* ctx, int n_plans, esl_plan* const* plans, const int* n_sets,
							   const double* const* d_rates, double* const* d_final, const esl_solve_opts* opts,
							   esl_solve_report* reports);

/* ---- device memory helpers (for callers without their own allocator) ------- */
esl_status esl_device_alloc(esl_ctx* ctx, size_t bytes, void** out);
esl_status esl_device_free(esl_ctx* ctx, void* p);
esl_status esl_memcpy_h2d(esl_ctx* ctx, void* dst, const void* src, size_t bytes);
esl_status esl_memcpy_d2h(esl_ctx* ctx, void* dst, const void* src, size_t bytes);
/* device to device on the ctx stream, by the library's own copy kernel (reads and
 * writes `bytes` once each: also a kernel of known traffic for counter scaling) */
esl_status esl_memcpy_d2d(esl_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---- kernel profiling (HIP events on the ctx stream) -----------------------
 * enable != 0 starts collecting per-kernel launch counts, device time and
 * algorithmic bytes (the bytes a launch must move at minimum, as defined next
 * to each kernel); enable == 0 stops.  Starting clears earlier figures. */
esl_status esl_profile(esl_ctx* ctx, int enable);
/* Figures of kernel `index` (0-based, in order of first launch); returns
 * ESL_E_ARG past the last one.  name_out receives up to name_cap bytes. */
esl_status esl_profile_stat(esl_ctx* ctx, int index, char* name_out, int name_cap, int64_t* launches,
							double* total_ms, double* total_bytes);

/* ---- stage timing (HIP events on the ctx stream) ---------------------------
 * Milliseconds of the device work of the last esl_table_build (stage 0) and
 * esl_graph_reorganize (stage 1) on g. */
esl_status esl_graph_stage_ms(const esl_graph* g, double ms[2]);

#ifdef __cplusplus
}
#endif

#endif /* ESL_H */
