// Graph stages on the device: SCC decomposition, condensation heights and the
// canonical vertex order.
//
// Replaces transition_graph::reorganize_vertices (reference
// cuda/src/transition_graph.cu:123-202), the GARDENIA FW-BW-trim SCC solver it
// calls (cuda/src/sga/two_phase.cu:12-145, bitset.cu, wcc.cu) and the
// topological_labelling / reorganize kernels (kernels/transition_graph.cu:3-52).
//
// The state graph is an orientation of (a subgraph of) the N-cube, so the
// graph is never stored as an adjacency list here: the neighbours of v are
// v ^ (1 << i), v -> v^(1<<i) iff bit i of mask[v], and v^(1<<i) -> v iff bit
// i of mask[v ^ (1 << i)].  One wavefront owns a 6-sub-cube (64 consecutive
// states), so the moves along bits 0..5 stay in registers (cross-lane
// shuffles) and each pass relaxes to a wave-local fixpoint before it touches
// HBM again; moves along bits >= 6 read the neighbouring sub-cube, which is
// a contiguous, coalesced 256-byte block.
//
// SCCs: trimming, then min-colour propagation along edges and a backward
// closure inside every colour class (the colouring SCC method).  Because a
// colour is the minimum vertex that reaches a vertex, the SCC that survives a
// round is named by its minimum vertex: the canonical SCC id, so the output is
// deterministic (the reference's atomicCAS pivots, bitset.cu:415-423, are not).
//
// Heights: round k labels every SCC whose successors are all labelled, which
// is the fixpoint of topological_labelling; terminal SCCs get 1.
//
// Order: [terminal SCCs by ascending id | transient SCCs by (height desc, id
// desc)], vertices ascending inside an SCC (transition_graph.cu:170-198).

#include <rocprim/device/device_radix_sort.hpp>

#include "common.hip.hpp"
#include "esl_internal.hpp"

namespace esl {

constexpr int gblock = 256;

// Moves along the high bits 6 .. N-1 read the neighbouring sub-cubes.  The
// kernels below take them in a fully unrolled loop of HMAX steps, step j on
// bit min(6 + j, N - 1) and counted only while 6 + j < N, so all of a lane's
// neighbour loads are in flight before the first is used (a loop bounded by
// the run-time N waits for each load in turn: one memory latency per bit).
// HMAX is the smallest of 8, 16, 26 that covers N - 6 (0 below N = 7).
__device__ __forceinline__ int high_bit(int j, int N) { return 6 + j < N ? 6 + j : N - 1; }

// Fixpoint passes are enqueued in groups without waiting for the host; a
// pass whose predecessor changed nothing (prev[0] == 0) returns at once.
__device__ __forceinline__ bool converged(const int32_t* prev) { return prev && prev[0] == 0; }

// The per-vertex bodies below are called by every lane of a wave together
// (they use cross-lane operations); v is the lane's vertex, valid when v < n.

// ---- trimming: a vertex without an active in- or out-neighbour is a singleton SCC
template <int HMAX>
__device__ __forceinline__ bool trim_vertex(int N, int64_t n, int64_t v, const uint32_t* __restrict__ mask,
											int32_t* __restrict__ scc)
{
	const bool valid = v < n;
	const uint32_t mv = valid ? mask[v] : 0u;
	bool active = valid && scc[v] < 0;
	bool gin = false, gout = false;
	if (HMAX > 0 && __any(active))
	{
		int32_t su[HMAX > 0 ? HMAX : 1];
		uint32_t mu[HMAX > 0 ? HMAX : 1];
		const int64_t vv = valid ? v : 0; // lanes past n read vertex 0's neighbours, unused
#pragma unroll
		for (int j = 0; j < HMAX; ++j)
		{
			const int64_t u = vv ^ (int64_t(1) << high_bit(j, N));
			su[j] = scc[u];
			mu[j] = mask[u];
		}
#pragma unroll
		for (int j = 0; j < HMAX; ++j)
		{
			const int i = 6 + j;
			const bool ua = i < N && su[j] < 0;
			gin |= ua && ((mu[j] >> i) & 1u);
			gout |= ua && ((mv >> i) & 1u);
		}
	}
	const int lim = N < 6 ? N : 6;
	bool trimmed = false;
	for (;;)
	{
		bool in = gin, out = gout;
		for (int i = 0; i < lim; ++i)
		{
			const int ua = __shfl_xor((int)active, 1 << i, wave);
			const uint32_t mu = __shfl_xor(mv, 1 << i, wave);
			in |= ua && ((mu >> i) & 1u);
			out |= ua && ((mv >> i) & 1u);
		}
		const bool t = active && !(in && out);
		if (!__any(t))
			break;
		if (t)
		{
			active = false;
			trimmed = true;
		}
	}
	if (trimmed)
		scc[v] = (int32_t)v;
	return trimmed;
}

template <int HMAX>
__global__ __launch_bounds__(gblock) void k_trim(int N, int64_t n, const uint32_t* __restrict__ mask,
												 int32_t* __restrict__ scc, int32_t* __restrict__ changed,
												 const int32_t* __restrict__ prev)
{
	if (converged(prev))
		return;
	if (trim_vertex<HMAX>(N, n, (int64_t)blockIdx.x * gblock + threadIdx.x, mask, scc))
		changed[0] = 1;
}

__global__ __launch_bounds__(gblock) void k_color_init(int64_t n, const int32_t* __restrict__ scc,
													   int32_t* __restrict__ col, uint8_t* __restrict__ inscc)
{
	const int64_t v = (int64_t)blockIdx.x * gblock + threadIdx.x;
	if (v < n)
	{
		col[v] = scc[v] < 0 ? (int32_t)v : INT32_MAX;
		inscc[v] = 0;
	}
}

// ---- min-colour propagation along edges (u -> v: col[v] = min(col[v], col[u]))
template <int HMAX>
__device__ __forceinline__ bool color_vertex(int N, int64_t n, int64_t v, const uint32_t* __restrict__ mask,
											 const int32_t* __restrict__ scc, int32_t* __restrict__ col)
{
	const bool valid = v < n;
	const bool active = valid && scc[v] < 0;
	const uint32_t mv = valid ? mask[v] : 0u;
	const int32_t c0 = active ? col[v] : INT32_MAX;
	int32_t c = c0;
	if (HMAX > 0 && __any(active))
	{
		uint32_t mu[HMAX > 0 ? HMAX : 1];
		int32_t su[HMAX > 0 ? HMAX : 1], cu[HMAX > 0 ? HMAX : 1];
		const int64_t vv = valid ? v : 0; // lanes past n read vertex 0's neighbours, unused
#pragma unroll
		for (int j = 0; j < HMAX; ++j)
		{
			const int64_t u = vv ^ (int64_t(1) << high_bit(j, N));
			mu[j] = mask[u];
			su[j] = scc[u];
			cu[j] = col[u];
		}
#pragma unroll
		for (int j = 0; j < HMAX; ++j)
		{
			const int i = 6 + j;
			if (active && i < N && ((mu[j] >> i) & 1u) && su[j] < 0)
				c = cu[j] < c ? cu[j] : c;
		}
	}
	const int lim = N < 6 ? N : 6;
	for (;;)
	{
		int32_t nc = c;
		for (int i = 0; i < lim; ++i)
		{
			const int32_t cu = __shfl_xor(c, 1 << i, wave);
			const uint32_t mu = __shfl_xor(mv, 1 << i, wave);
			if (active && ((mu >> i) & 1u))
				nc = cu < nc ? cu : nc;
		}
		const bool ch = nc != c;
		c = nc;
		if (!__any(ch))
			break;
	}
	if (active && c != c0)
	{
		col[v] = c;
		return true;
	}
	return false;
}

template <int HMAX>
__global__ __launch_bounds__(gblock) void k_color(int N, int64_t n, const uint32_t* __restrict__ mask,
												  const int32_t* __restrict__ scc, int32_t* __restrict__ col,
												  int32_t* __restrict__ changed, const int32_t* __restrict__ prev)
{
	if (converged(prev))
		return;
	if (color_vertex<HMAX>(N, n, (int64_t)blockIdx.x * gblock + threadIdx.x, mask, scc, col))
		changed[0] = 1;
}

// ---- backward closure inside a colour class: v joins the SCC of its colour's
// root if it has an edge into a member with the same colour.
template <int HMAX>
__device__ __forceinline__ bool close_vertex(int N, int64_t n, int64_t v, const uint32_t* __restrict__ mask,
											 const int32_t* __restrict__ scc, const int32_t* __restrict__ col,
											 uint8_t* __restrict__ inscc)
{
	const bool valid = v < n;
	const bool active = valid && scc[v] < 0;
	const uint32_t mv = valid ? mask[v] : 0u;
	const int32_t c = active ? col[v] : -1;
	const bool in0 = active && inscc[v];
	bool in = in0 || (active && c == (int32_t)v); // the colour's root is in its own SCC
	if (HMAX > 0 && __any(active && !in))
	{
		uint8_t iw[HMAX > 0 ? HMAX : 1];
		int32_t cw[HMAX > 0 ? HMAX : 1], sw[HMAX > 0 ? HMAX : 1];
		const int64_t vv = valid ? v : 0; // lanes past n read vertex 0's neighbours, unused
#pragma unroll
		for (int j = 0; j < HMAX; ++j)
		{
			const int64_t w = vv ^ (int64_t(1) << high_bit(j, N));
			iw[j] = inscc[w];
			cw[j] = col[w];
			sw[j] = scc[w];
		}
		bool hit = false;
#pragma unroll
		for (int j = 0; j < HMAX; ++j)
		{
			const int i = 6 + j;
			hit |= i < N && ((mv >> i) & 1u) && iw[j] && cw[j] == c && sw[j] < 0;
		}
		if (active && hit)
			in = true;
	}
	const int lim = N < 6 ? N : 6;
	for (;;)
	{
		bool nin = in;
		for (int i = 0; i < lim; ++i)
		{
			const int wi = __shfl_xor((int)in, 1 << i, wave);
			const int32_t cw = __shfl_xor(c, 1 << i, wave);
			if (active && ((mv >> i) & 1u) && wi && cw == c)
				nin = true;
		}
		const bool ch = nin != in;
		in = nin;
		if (!__any(ch))
			break;
	}
	if (in && !in0)
	{
		inscc[v] = 1;
		return true;
	}
	return false;
}

template <int HMAX>
__global__ __launch_bounds__(gblock) void k_close(int N, int64_t n, const uint32_t* __restrict__ mask,
												  const int32_t* __restrict__ scc, const int32_t* __restrict__ col,
												  uint8_t* __restrict__ inscc, int32_t* __restrict__ changed,
												  const int32_t* __restrict__ prev)
{
	if (converged(prev))
		return;
	if (close_vertex<HMAX>(N, n, (int64_t)blockIdx.x * gblock + threadIdx.x, mask, scc, col, inscc))
		changed[0] = 1;
}

__global__ __launch_bounds__(gblock) void k_assign(int64_t n, int32_t* __restrict__ scc,
												   const int32_t* __restrict__ col, const uint8_t* __restrict__ inscc,
												   int32_t* __restrict__ remaining)
{
	const int64_t v = (int64_t)blockIdx.x * gblock + threadIdx.x;
	int left = 0;
	if (v < n && scc[v] < 0)
	{
		if (inscc[v])
			scc[v] = col[v];
		else
			left = 1;
	}
	const int cnt = __popcll(__ballot(left));
	if (lane_id() == 0 && cnt)
		atomicAdd(remaining, cnt);
}

// ---- heights.  Level `cur` is enqueued in groups like the fixpoint passes;
// once every root is labelled (total[0] == S) the kernels return at once.
template <int HMAX>
__device__ __forceinline__ void height_check_vertex(int N, int64_t n, int64_t v, const uint32_t* __restrict__ mask,
													const int32_t* __restrict__ scc, const int32_t* __restrict__ h,
													int32_t* __restrict__ blocked, int32_t cur)
{
	if (v >= n)
		return;
	const int32_t r = scc[v];
	if (h[r] != 0)
		return;
	const uint32_t mv = mask[v];
	constexpr int B = HMAX + 6; // bits 0 .. N-1, clamped like the high bits
	int32_t rw[B];
#pragma unroll
	for (int j = 0; j < B; ++j)
		rw[j] = scc[v ^ (int64_t(1) << (j < N ? j : N - 1))];
	int32_t hw[B];
#pragma unroll
	for (int j = 0; j < B; ++j)
		hw[j] = h[rw[j]]; // every lane gathers, so the loads are issued together
	bool blk = false;
#pragma unroll
	for (int j = 0; j < B; ++j)
		blk |= j < N && ((mv >> j) & 1u) && rw[j] != r && hw[j] == 0;
	if (blk)
		blocked[r] = cur;
}

// 1 when v is a root still unlabelled and not blocked at level cur (then labelled cur)
__device__ __forceinline__ int height_assign_vertex(int64_t n, int64_t v, const int32_t* __restrict__ scc,
													int32_t* __restrict__ h, const int32_t* __restrict__ blocked,
													int32_t cur)
{
	if (v < n && scc[v] == (int32_t)v && h[v] == 0 && blocked[v] != cur)
	{
		h[v] = cur;
		return 1;
	}
	return 0;
}

template <int HMAX>
__global__ __launch_bounds__(gblock) void k_height_check(int N, int64_t n, const uint32_t* __restrict__ mask,
														 const int32_t* __restrict__ scc,
														 const int32_t* __restrict__ h, int32_t* __restrict__ blocked,
														 int32_t cur, const int32_t* __restrict__ total, int32_t S)
{
	if (total[0] >= S)
		return;
	height_check_vertex<HMAX>(N, n, (int64_t)blockIdx.x * gblock + threadIdx.x, mask, scc, h, blocked, cur);
}

__global__ __launch_bounds__(gblock) void k_height_assign(int64_t n, const int32_t* __restrict__ scc,
														  int32_t* __restrict__ h,
														  const int32_t* __restrict__ blocked, int32_t cur,
														  int32_t* __restrict__ labelled, int32_t* __restrict__ total,
														  int32_t S)
{
	if (total[0] >= S)
		return;
	const int64_t v = (int64_t)blockIdx.x * gblock + threadIdx.x;
	const int got = height_assign_vertex(n, v, scc, h, blocked, cur);
	const int w = __popcll(__ballot(got));
	if (lane_id() == 0 && w)
		atomicAdd(labelled, w);
}

// the labelled count of a level into the running total (one thread, after the level)
__global__ void k_height_total(const int32_t* __restrict__ labelled, int32_t* __restrict__ total)
{
	total[0] += labelled[0];
}

// ---- canonical order
// root key: terminal -> (0, id); transient -> (H + 1 - h, n - 1 - id)
__global__ __launch_bounds__(gblock) void k_root_keys(int64_t n, const int32_t* __restrict__ scc,
													  const int32_t* __restrict__ h, int32_t H,
													  const int32_t* __restrict__ root_pos,
													  uint64_t* __restrict__ keys, int32_t* __restrict__ vals)
{
	const int64_t v = (int64_t)blockIdx.x * gblock + threadIdx.x;
	if (v >= n)
		return;
	const int32_t r = scc[v];
	if (r != (int32_t)v)
		return;
	const int32_t hv = h[v];
	const uint64_t hi = hv == 1 ? 0ull : (uint64_t)(H + 1 - hv);
	const uint64_t lo = hv == 1 ? (uint64_t)v : (uint64_t)(n - 1 - v);
	const int32_t p = root_pos[v];
	keys[p] = (hi << 32) | lo;
	vals[p] = (int32_t)v;
}

__global__ __launch_bounds__(gblock) void k_is_root(int64_t n, const int32_t* __restrict__ scc,
													int32_t* __restrict__ flag)
{
	const int64_t v = (int64_t)blockIdx.x * gblock + threadIdx.x;
	if (v < n)
		flag[v] = scc[v] == (int32_t)v ? 1 : 0;
}

__global__ __launch_bounds__(gblock) void k_rank_roots(int64_t S, const int32_t* __restrict__ sorted_roots,
													   int32_t* __restrict__ rank)
{
	const int64_t k = (int64_t)blockIdx.x * gblock + threadIdx.x;
	if (k >= S)
		return;
	rank[sorted_roots[k]] = (int32_t)k; // rank is indexed by root vertex here
}

// sccs_offsets from the vertices sorted by SCC rank: offsets[k] = first position of rank k
__global__ __launch_bounds__(gblock) void k_seg_bounds(int64_t n, int64_t S, const uint32_t* __restrict__ sorted_keys,
													   int32_t* __restrict__ offsets)
{
	const int64_t p = (int64_t)blockIdx.x * gblock + threadIdx.x;
	if (p >= n)
		return;
	const uint32_t k = sorted_keys[p];
	if (p == 0 || sorted_keys[p - 1] != k)
		offsets[k] = (int32_t)p;
	if (p == n - 1)
		offsets[S] = (int32_t)n;
}

__global__ __launch_bounds__(gblock) void k_vertex_keys(int64_t n, const int32_t* __restrict__ scc,
														const int32_t* __restrict__ h,
														const int32_t* __restrict__ root_rank,
														uint32_t* __restrict__ keys, int32_t* __restrict__ vals,
														int32_t* __restrict__ vrank, int32_t* __restrict__ vheight)
{
	const int64_t v = (int64_t)blockIdx.x * gblock + threadIdx.x;
	if (v >= n)
		return;
	const int32_t r = scc[v];
	const int32_t k = root_rank[r];
	keys[v] = (uint32_t)k;
	vals[v] = (int32_t)v;
	vrank[v] = k;
	vheight[v] = h[r];
}

namespace {

struct pass_flags
{
	ctx_t& c;
	int32_t* d;
	explicit pass_flags(ctx_t& ctx) : c(ctx), d(ctx.flags.get()) {}
	void clear(int k) { ESL_HIP(hipMemsetAsync(d + k, 0, sizeof(int32_t), c.stream)); }
	int32_t read(int k)
	{
		ESL_HIP(hipMemcpyAsync(c.h_flags + k, d + k, sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
		ESL_HIP(hipStreamSynchronize(c.stream));
		return c.h_flags[k];
	}
};

// Runs a fixpoint pass in groups of launches with one host round trip per
// group: pass p of a group reports a change in its own flag slot; the first pass
// that changed nothing ends the loop (the passes after it in its group return
// at once, seeing their predecessor's flag clear).  pass(flag, prev) launches
// one pass (prev: the previous pass's flag, null for a group's first).  The
// first group has `hint` passes (what the same fixpoint needed last time, plus
// one), later groups `group`; hint is updated to the passes used plus one.
template <typename F>
void fixpoint(ctx_t& c, graph_t& g, int& hint, int group, int max_passes, const char* what, F&& pass)
{
	constexpr int base = 8, slots = 16; // flag slots base .. base + slots - 1
	int32_t* d = c.flags.get() + base;
	for (int done = 0;;)
	{
		++g.host_syncs;
		const int gsz = std::min(slots, std::max(2, done ? group : hint));
		ESL_HIP(hipMemsetAsync(d, 0, gsz * sizeof(int32_t), c.stream));
		for (int p = 0; p < gsz; ++p)
			pass(d + p, p ? d + p - 1 : nullptr);
		ESL_HIP(hipMemcpyAsync(c.h_flags + base, d, gsz * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
		ESL_HIP(hipStreamSynchronize(c.stream));
		for (int p = 0; p < gsz; ++p)
			if (!c.h_flags[base + p])
			{
				hint = done + p + 2;
				g.mask_passes += done + p + 1;
				return;
			}
		done += gsz;
		if (done > max_passes)
			throw error(ESL_E_DEVICE, std::string(what) + " did not converge");
	}
}

// the unrolled width of the high-bit loops for N state bits
int hmax_for(int N) { return N <= 6 ? 0 : N <= 14 ? 8 : N <= 22 ? 16 : 26; }

#define ESL_HMAX_DISPATCH(N, KERNEL, ...)                                                                              \
	switch (hmax_for(N))                                                                                               \
	{                                                                                                                  \
	case 0:                                                                                                            \
		KERNEL<0>__VA_ARGS__;                                                                                          \
		break;                                                                                                         \
	case 8:                                                                                                            \
		KERNEL<8>__VA_ARGS__;                                                                                          \
		break;                                                                                                         \
	case 16:                                                                                                           \
		KERNEL<16>__VA_ARGS__;                                                                                         \
		break;                                                                                                         \
	default:                                                                                                           \
		KERNEL<26>__VA_ARGS__;                                                                                         \
		break;                                                                                                         \
	}

} // namespace

void graph_reorganize(ctx_t& c, graph_t& g)
{
	const int N = g.N;
	const int64_t n = g.n;
	const int nb = ceil_div(n, gblock);
	pass_flags fl(c);

	g.scc.ensure(n);
	dbuf<int32_t> col;
	col.ensure(n);
	dbuf<uint8_t> inscc;
	inscc.ensure(n);

	ESL_HIP(hipEventRecord(c.ev[2], c.stream));
	ESL_HIP(hipMemsetAsync(g.scc.get(), 0xff, n * sizeof(int32_t), c.stream));

	const int max_passes = 1 << 20;
	const int pass_group = 8;	   // fixpoint passes per host round trip, after the first
	int hint[3] = { 8, 8, 8 }; // first-group sizes of the three fixpoints
	int64_t prev_left = n + 1;
	g.mask_passes = g.scc_rounds = g.host_syncs = 0;
	for (int round = 0;; ++round)
	{
		++g.scc_rounds;
		// algorithmic bytes of one pass over the flip masks: mask (4 B/state) and the
		// state labels it reads and writes
		fixpoint(c, g, hint[0], pass_group, max_passes, "trim", [&](int32_t* f, const int32_t* pf) {
			timed_launch(c, "k_trim", 12.0 * n, [&] {
				ESL_HMAX_DISPATCH(N, k_trim, <<<nb, gblock, 0, c.stream>>>(N, n, g.mask.get(), g.scc.get(), f, pf));
			});
		});
		k_color_init<<<nb, gblock, 0, c.stream>>>(n, g.scc.get(), col.get(), inscc.get());
		fixpoint(c, g, hint[1], pass_group, max_passes, "colouring", [&](int32_t* f, const int32_t* pf) {
			timed_launch(c, "k_color", 16.0 * n, [&] {
				ESL_HMAX_DISPATCH(N, k_color,
								  <<<nb, gblock, 0, c.stream>>>(N, n, g.mask.get(), g.scc.get(), col.get(), f, pf));
			});
		});
		fixpoint(c, g, hint[2], pass_group, max_passes, "closure", [&](int32_t* f, const int32_t* pf) {
			timed_launch(c, "k_close", 14.0 * n, [&] {
				ESL_HMAX_DISPATCH(N, k_close,
								  <<<nb, gblock, 0, c.stream>>>(N, n, g.mask.get(), g.scc.get(), col.get(),
																inscc.get(), f, pf));
			});
		});
		fl.clear(1);
		timed_launch(c, "k_assign", 13.0 * n,
					 [&] { k_assign<<<nb, gblock, 0, c.stream>>>(n, g.scc.get(), col.get(), inscc.get(), fl.d + 1); });
		ESL_HIP(hipGetLastError());
		const int32_t left = fl.read(1);
		++g.host_syncs;
		if (!left)
			break;
		// every round removes at least the SCC of the smallest active vertex
		if (left >= prev_left)
			throw error(ESL_E_DEVICE, "SCC round made no progress");
		prev_left = left;
	}

	// heights (per root), counted until every root is labelled
	dbuf<int32_t> is_root, root_pos;
	is_root.ensure(n);
	root_pos.ensure(n + 1);
	k_is_root<<<nb, gblock, 0, c.stream>>>(n, g.scc.get(), is_root.get());
	exclusive_scan_i32(c, is_root.get(), root_pos.get(), n);
	int32_t S = 0;
	ESL_HIP(hipMemcpyAsync(&S, root_pos.get() + n, sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
	ESL_HIP(hipStreamSynchronize(c.stream));
	g.S = S;

	dbuf<int32_t> hroot, blocked;
	hroot.ensure(n);
	blocked.ensure(n);
	ESL_HIP(hipMemsetAsync(hroot.get(), 0, n * sizeof(int32_t), c.stream));
	ESL_HIP(hipMemsetAsync(blocked.get(), 0, n * sizeof(int32_t), c.stream));
	// levels in groups of `level_group` per host round trip: slot hbase + k counts
	// the roots labelled at the group's k-th level, slot htotal all so far
	constexpr int level_group = 16, hbase = 32, htotal = 31;
	int32_t* dl = c.flags.get() + hbase;
	int32_t* dtot = c.flags.get() + htotal;
	ESL_HIP(hipMemsetAsync(dtot, 0, sizeof(int32_t), c.stream));
	int32_t labelled_total = 0, T = 0, cur = 0, H = 0;
	while (!H)
	{
		ESL_HIP(hipMemsetAsync(dl, 0, level_group * sizeof(int32_t), c.stream));
		for (int k = 0; k < level_group; ++k)
		{
			const int32_t lev = cur + 1 + k;
			timed_launch(c, "k_height_check", 13.0 * n, [&] {
				ESL_HMAX_DISPATCH(N, k_height_check,
								  <<<nb, gblock, 0, c.stream>>>(N, n, g.mask.get(), g.scc.get(), hroot.get(),
																blocked.get(), lev, dtot, S));
			});
			timed_launch(c, "k_height_assign", 9.0 * n, [&] {
				k_height_assign<<<nb, gblock, 0, c.stream>>>(n, g.scc.get(), hroot.get(), blocked.get(), lev, dl + k,
															 dtot, S);
			});
			k_height_total<<<1, 1, 0, c.stream>>>(dl + k, dtot);
		}
		ESL_HIP(hipGetLastError());
		ESL_HIP(hipMemcpyAsync(c.h_flags + hbase, dl, level_group * sizeof(int32_t), hipMemcpyDeviceToHost,
							   c.stream));
		ESL_HIP(hipStreamSynchronize(c.stream));
		++g.host_syncs;
		for (int k = 0; k < level_group && !H; ++k)
		{
			const int32_t got = c.h_flags[hbase + k];
			++g.mask_passes;
			if (got == 0)
				throw error(ESL_E_DEVICE, "condensation graph is not acyclic");
			if (cur + 1 + k == 1)
				T = got;
			labelled_total += got;
			if (labelled_total >= S)
				H = cur + 1 + k;
		}
		cur += level_group;
	}
	g.T = T;
	g.H = H;

	// sort roots by key
	dbuf<uint64_t> rkeys, rkeys_out;
	dbuf<int32_t> rvals, rvals_out;
	rkeys.ensure(S);
	rkeys_out.ensure(S);
	rvals.ensure(S);
	rvals_out.ensure(S);
	k_root_keys<<<nb, gblock, 0, c.stream>>>(n, g.scc.get(), hroot.get(), g.H, root_pos.get(), rkeys.get(),
											 rvals.get());
	ESL_HIP(hipGetLastError());
	size_t tmp_bytes = 0;
	ESL_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, rkeys.get(), rkeys_out.get(), rvals.get(),
									  rvals_out.get(), (size_t)S, 0, 64, c.stream));
	dbuf<uint8_t> tmp;
	tmp.ensure(tmp_bytes);
	ESL_HIP(rocprim::radix_sort_pairs(tmp.get(), tmp_bytes, rkeys.get(), rkeys_out.get(), rvals.get(),
									  rvals_out.get(), (size_t)S, 0, 64, c.stream));
	dbuf<int32_t> root_rank;
	root_rank.ensure(n);
	k_rank_roots<<<ceil_div(S, gblock), gblock, 0, c.stream>>>(S, rvals_out.get(), root_rank.get());

	// stable sort of the vertices by their SCC's rank
	dbuf<uint32_t> vkeys, vkeys_out;
	dbuf<int32_t> vvals;
	vkeys.ensure(n);
	vkeys_out.ensure(n);
	vvals.ensure(n);
	g.rank.ensure(n);
	g.height.ensure(n);
	g.reordered.ensure(n);
	k_vertex_keys<<<nb, gblock, 0, c.stream>>>(n, g.scc.get(), hroot.get(), root_rank.get(), vkeys.get(),
											   vvals.get(), g.rank.get(), g.height.get());
	ESL_HIP(hipGetLastError());
	unsigned end_bit = 1;
	while ((int64_t(1) << end_bit) < (int64_t)S)
		++end_bit;
	tmp_bytes = 0;
	ESL_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, vkeys.get(), vkeys_out.get(), vvals.get(),
									  g.reordered.get(), (size_t)n, 0, end_bit, c.stream));
	tmp.ensure(tmp_bytes);
	ESL_HIP(rocprim::radix_sort_pairs(tmp.get(), tmp_bytes, vkeys.get(), vkeys_out.get(), vvals.get(),
									  g.reordered.get(), (size_t)n, 0, end_bit, c.stream));
	g.offsets.ensure(S + 1);
	k_seg_bounds<<<nb, gblock, 0, c.stream>>>(n, S, vkeys_out.get(), g.offsets.get());
	ESL_HIP(hipGetLastError());
	ESL_HIP(hipEventRecord(c.ev[3], c.stream));
	ESL_HIP(hipEventSynchronize(c.ev[3]));
	float ms = 0;
	ESL_HIP(hipEventElapsedTime(&ms, c.ev[2], c.ev[3]));
	g.stage_ms[1] = ms;

	g.h_offsets.resize(S + 1);
	ESL_HIP(hipMemcpyAsync(g.h_offsets.data(), g.offsets.get(), (S + 1) * sizeof(int32_t), hipMemcpyDeviceToHost,
						   c.stream));
	ESL_HIP(hipStreamSynchronize(c.stream));
	g.n_t = g.h_offsets[T];
	g.reorganized = true;
}

// ---- graph stage results supplied by the caller (a stored .symb file)

// rank[v] = index of v's SCC in the canonical order, scc[v] = its minimum vertex
// (the first of the SCC's block, vertices being ascending inside a block)
__global__ __launch_bounds__(gblock) void k_import_rank(int64_t n, int64_t S, const int32_t* __restrict__ offsets,
														const int32_t* __restrict__ reordered, int32_t* __restrict__ rank,
														int32_t* __restrict__ scc)
{
	const int64_t p = (int64_t)blockIdx.x * gblock + threadIdx.x;
	if (p >= n)
		return;
	int64_t lo = 0, hi = S; // offsets[lo] <= p < offsets[hi]
	while (hi - lo > 1)
	{
		const int64_t mid = (lo + hi) / 2;
		if (offsets[mid] <= p)
			lo = mid;
		else
			hi = mid;
	}
	const int32_t v = reordered[p];
	rank[v] = (int32_t)lo;
	scc[v] = reordered[offsets[lo]];
}

void graph_import(ctx_t& c, graph_t& g, const int32_t* h_reordered, const int32_t* h_offsets, int64_t S, int64_t T)
{
	const int64_t n = g.n;
	if (S < 1 || T < 1 || T > S || S > n)
		throw error(ESL_E_ARG, "graph import: bad SCC counts");
	if (h_offsets[0] != 0 || h_offsets[S] != n)
		throw error(ESL_E_ARG, "graph import: offsets must start at 0 and end at the state count");
	for (int64_t k = 0; k < S; ++k)
		if (h_offsets[k + 1] <= h_offsets[k])
			throw error(ESL_E_ARG, "graph import: empty or decreasing SCC block");
	std::vector<uint8_t> seen(n, 0);
	for (int64_t k = 0; k < S; ++k)
		for (int32_t p = h_offsets[k]; p < h_offsets[k + 1]; ++p)
		{
			const int32_t v = h_reordered[p];
			if (v < 0 || v >= n || seen[v])
				throw error(ESL_E_ARG, "graph import: reordered vertices are not a permutation");
			seen[v] = 1;
			if (p > h_offsets[k] && v <= h_reordered[p - 1])
				throw error(ESL_E_ARG, "graph import: vertices must ascend inside an SCC");
		}
	g.reordered.ensure(n);
	g.offsets.ensure(S + 1);
	g.rank.ensure(n);
	g.scc.ensure(n);
	ESL_HIP(hipMemcpyAsync(g.reordered.get(), h_reordered, n * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
	ESL_HIP(hipMemcpyAsync(g.offsets.get(), h_offsets, (S + 1) * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
	k_import_rank<<<ceil_div(n, gblock), gblock, 0, c.stream>>>(n, S, g.offsets.get(), g.reordered.get(),
																g.rank.get(), g.scc.get());
	ESL_HIP(hipGetLastError());
	ESL_HIP(hipStreamSynchronize(c.stream));
	g.h_offsets.assign(h_offsets, h_offsets + S + 1);
	g.S = S;
	g.T = T;
	g.n_t = h_offsets[T];
	g.H = 0; // heights are not stored
	g.stage_ms[1] = 0.0;
	g.sol = solution_t{};
	g.reorganized = true;
}

} // namespace esl
