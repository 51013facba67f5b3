// Transition table on the device.
//
// Replaces transition_table::construct_table (reference
// cuda/src/transition_table.cu:76-174): there, every node costs (clauses + 2)
// full passes over the 2^N states plus a global COO sort by source
// (cusparseXcoosortByColumn).  Here one thread owns one state s, evaluates
// every rule from its truth table (held in LDS), and produces the flip mask
//     mask[s] bit i  <=>  f_i(s) != s_i  and the rate of that direction > 0,
// which is the whole graph in 4 bytes per state: the successors of s are
// s ^ (1 << i) for the set bits, in ascending node order.  The CSC arrays the
// reference exposes (indptr, rows = target, cols = source) are then written
// from the mask: a workgroup scans its degrees in LDS, stages its rows/cols in
// LDS and writes them to HBM as one contiguous, coalesced run.

#include "common.hip.hpp"
#include "esl_internal.hpp"

namespace esl {

constexpr int table_block = 256;
constexpr int lds_table_words = 8192; // 32 KiB of truth-table bits staged in LDS

__global__ __launch_bounds__(table_block) void k_flip_mask(int N, int64_t n, const rule_dev_t* __restrict__ rules,
														   const int32_t* __restrict__ inputs,
														   const uint32_t* __restrict__ tables, int table_words,
														   uint32_t up_ok, uint32_t down_ok,
														   uint32_t* __restrict__ mask, int32_t* __restrict__ block_deg)
{
	__shared__ uint32_t s_tab[lds_table_words];
	__shared__ int32_t s_red[table_block / wave + 1];
	const bool tab_in_lds = table_words <= lds_table_words;
	if (tab_in_lds)
		for (int i = threadIdx.x; i < table_words; i += table_block)
			s_tab[i] = tables[i];
	__syncthreads();

	const int64_t s = (int64_t)blockIdx.x * table_block + threadIdx.x;
	int deg = 0;
	if (s < n)
	{
		const uint32_t st = (uint32_t)s;
		uint32_t m = 0;
		for (int i = 0; i < N; ++i)
		{
			const rule_dev_t r = rules[i];
			uint32_t idx = 0;
			for (int k = 0; k < r.k; ++k)
				idx |= ((st >> inputs[r.in_off + k]) & 1u) << k;
			const uint32_t bit = r.tab_off + idx;
			const uint32_t word = tab_in_lds ? s_tab[bit >> 5] : tables[bit >> 5];
			const uint32_t f = (word >> (bit & 31)) & 1u;
			const uint32_t cur = (st >> i) & 1u;
			const uint32_t allowed = cur ? ((down_ok >> i) & 1u) : ((up_ok >> i) & 1u);
			m |= ((f ^ cur) & allowed) << i;
		}
		mask[s] = m;
		deg = __popc(m);
	}
	const int tot = block_sum<table_block>(deg, s_red);
	if (threadIdx.x == 0)
		block_deg[blockIdx.x] = tot;
}

// The rules of a network whose every node reads at most 6 nodes, by value in
// the kernel arguments (uniform, so the loads are scalar): per node its input
// indices packed 5 bits each, its truth table as one 64-bit word and its
// input count.  k_flip_mask_small evaluates a rule with a few shifts and one
// table lookup in registers, where k_flip_mask reads the rule, then its inputs,
// then the table word from memory for every node of every state.
constexpr int small_rule_inputs = 6;
struct rules_small_t
{
	uint64_t tab[32];
	uint32_t pin[32];
	uint8_t k[32];
};

__global__ __launch_bounds__(table_block) void k_flip_mask_small(int N, int64_t n, rules_small_t R, uint32_t up_ok,
																 uint32_t down_ok, uint32_t* __restrict__ mask,
																 int32_t* __restrict__ block_deg)
{
	__shared__ int32_t s_red[table_block / wave + 1];
	const int64_t s = (int64_t)blockIdx.x * table_block + threadIdx.x;
	int deg = 0;
	if (s < n)
	{
		const uint32_t st = (uint32_t)s;
		uint32_t m = 0;
		for (int i = 0; i < N; ++i)
		{
			const uint32_t p = R.pin[i];
			const int k = R.k[i];
			uint32_t idx = 0;
#pragma unroll
			for (int q = 0; q < small_rule_inputs; ++q)
				if (q < k)
					idx |= ((st >> ((p >> (5 * q)) & 31u)) & 1u) << q;
			const uint32_t f = (uint32_t)(R.tab[i] >> idx) & 1u;
			const uint32_t cur = (st >> i) & 1u;
			const uint32_t allowed = cur ? ((down_ok >> i) & 1u) : ((up_ok >> i) & 1u);
			m |= ((f ^ cur) & allowed) << i;
		}
		mask[s] = m;
		deg = __popc(m);
	}
	const int tot = block_sum<table_block>(deg, s_red);
	if (threadIdx.x == 0)
		block_deg[blockIdx.x] = tot;
}

// indptr[s] = block offset + in-block exclusive prefix of the degrees; rows and
// cols are staged in LDS and written as the block's contiguous output run.
__global__ __launch_bounds__(table_block) void k_table_fill(int N, int64_t n, const uint32_t* __restrict__ mask,
															const int32_t* __restrict__ block_off,
															int32_t* __restrict__ indptr, int32_t* __restrict__ rows,
															int32_t* __restrict__ cols)
{
	extern __shared__ int32_t s_dyn[];
	__shared__ int32_t s_red[table_block / wave + 1];
	int32_t* s_rows = s_dyn;
	int32_t* s_cols = s_dyn + table_block * N;

	const int64_t s = (int64_t)blockIdx.x * table_block + threadIdx.x;
	const uint32_t m = s < n ? mask[s] : 0u;
	const int deg = __popc(m);
	int total;
	const int local = block_excl_scan<table_block>(deg, s_red, total);
	const int32_t base = block_off[blockIdx.x];
	if (s < n)
	{
		indptr[s] = base + local;
		if (s == n - 1)
			indptr[n] = base + local + deg;
		uint32_t rest = m;
		int o = local;
		while (rest)
		{
			const int i = __ffs(rest) - 1;
			rest &= rest - 1;
			s_rows[o] = (int32_t)((uint32_t)s ^ (1u << i));
			s_cols[o] = (int32_t)s;
			++o;
		}
	}
	__syncthreads();
	for (int i = threadIdx.x; i < total; i += table_block)
	{
		rows[base + i] = s_rows[i];
		cols[base + i] = s_cols[i];
	}
}

// ---- generic device-wide exclusive scan (int32) ------------------------------

constexpr int scan_block = 256;
constexpr int scan_items = 4;
constexpr int scan_tile = scan_block * scan_items;

__global__ __launch_bounds__(scan_block) void k_scan_reduce(const int32_t* __restrict__ in, int64_t n,
															int32_t* __restrict__ sums)
{
	__shared__ int32_t s_red[scan_block / wave + 1];
	const int64_t base = (int64_t)blockIdx.x * scan_tile;
	int v = 0;
	for (int k = 0; k < scan_items; ++k)
	{
		const int64_t i = base + (int64_t)k * scan_block + threadIdx.x;
		if (i < n)
			v += in[i];
	}
	v = block_sum<scan_block>(v, s_red);
	if (threadIdx.x == 0)
		sums[blockIdx.x] = v;
}

// Single workgroup: exclusive scan of `count` values in place, total appended.
__global__ __launch_bounds__(1024) void k_scan_single(int32_t* __restrict__ v, int64_t count)
{
	__shared__ int32_t s_red[1024 / wave + 1];
	int32_t carry = 0;
	for (int64_t base = 0; base < count; base += 1024)
	{
		const int64_t i = base + threadIdx.x;
		const int x = i < count ? v[i] : 0;
		int total;
		const int ex = block_excl_scan<1024>(x, s_red, total);
		if (i < count)
			v[i] = carry + ex;
		carry += total;
	}
	if (threadIdx.x == 0)
		v[count] = carry;
}

__global__ __launch_bounds__(scan_block) void k_scan_down(const int32_t* __restrict__ in, int64_t n,
														  const int32_t* __restrict__ sums, int32_t* __restrict__ out)
{
	__shared__ int32_t s_red[scan_block / wave + 1];
	const int64_t base = (int64_t)blockIdx.x * scan_tile;
	int32_t carry = sums[blockIdx.x];
	for (int k = 0; k < scan_items; ++k)
	{
		const int64_t i = base + (int64_t)k * scan_block + threadIdx.x;
		const int x = i < n ? in[i] : 0;
		int total;
		const int ex = block_excl_scan<scan_block>(x, s_red, total);
		if (i < n)
			out[i] = carry + ex;
		carry += total;
	}
	if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
		out[n] = carry;
}

void exclusive_scan_i32(ctx_t& c, const int32_t* in, int32_t* out, int64_t n)
{
	const int nb = ceil_div(n, scan_tile);
	dbuf<int32_t> sums;
	sums.ensure(nb + 1);
	k_scan_reduce<<<nb, scan_block, 0, c.stream>>>(in, n, sums.get());
	k_scan_single<<<1, 1024, 0, c.stream>>>(sums.get(), nb);
	k_scan_down<<<nb, scan_block, 0, c.stream>>>(in, n, sums.get(), out);
	ESL_HIP(hipGetLastError());
	ESL_HIP(hipStreamSynchronize(c.stream)); // `sums` is freed on return
}

// ---- host side -----------------------------------------------------------------

void table_build(ctx_t& c, graph_t& g, const model_t& m)
{
	const int N = (int)m.nodes.size();
	g.N = N;
	g.n = int64_t(1) << N;

	std::vector<rule_dev_t> rd(N);
	std::vector<int32_t> ins;
	std::vector<uint32_t> tabs;
	int64_t bits = 0;
	for (int i = 0; i < N; ++i)
	{
		const rule_t& r = m.rules[i];
		rd[i].k = (int32_t)r.inputs.size();
		rd[i].in_off = (int32_t)ins.size();
		rd[i].tab_off = (int32_t)bits;
		rd[i].pad = 0;
		ins.insert(ins.end(), r.inputs.begin(), r.inputs.end());
		bits += (int64_t)r.table.size();
	}
	if (bits > (int64_t(1) << 31))
		throw error(ESL_E_ARG, "rule truth tables too large");
	tabs.assign((size_t)((bits + 31) / 32), 0u);
	for (int i = 0; i < N; ++i)
	{
		const rule_t& r = m.rules[i];
		for (size_t k = 0; k < r.table.size(); ++k)
			if (r.table[k])
			{
				const int64_t b = rd[i].tab_off + (int64_t)k;
				tabs[b >> 5] |= 1u << (b & 31);
			}
	}
	if (ins.empty())
		ins.push_back(0);
	g.rules.ensure(N);
	g.rule_inputs.ensure(ins.size());
	g.rule_tables.ensure(tabs.size());
	ESL_HIP(hipMemcpyAsync(g.rules.get(), rd.data(), N * sizeof(rule_dev_t), hipMemcpyHostToDevice, c.stream));
	ESL_HIP(hipMemcpyAsync(g.rule_inputs.get(), ins.data(), ins.size() * sizeof(int32_t), hipMemcpyHostToDevice,
						   c.stream));
	ESL_HIP(hipMemcpyAsync(g.rule_tables.get(), tabs.data(), tabs.size() * sizeof(uint32_t),
						   hipMemcpyHostToDevice, c.stream));

	const int64_t n = g.n;
	const int nb = ceil_div(n, table_block);
	g.mask.ensure(n);
	g.indptr.ensure(n + 1);
	dbuf<int32_t> block_deg;
	block_deg.ensure(nb + 1);

	bool small = N <= 32;
	for (int i = 0; i < N && small; ++i)
		small = (int)m.rules[i].inputs.size() <= small_rule_inputs;
	ESL_HIP(hipEventRecord(c.ev[0], c.stream));
	// algorithmic bytes: the 4-byte mask written per state
	if (small)
	{
		rules_small_t R{};
		for (int i = 0; i < N; ++i)
		{
			const rule_t& r = m.rules[i];
			R.k[i] = (uint8_t)r.inputs.size();
			for (size_t q = 0; q < r.inputs.size(); ++q)
				R.pin[i] |= (uint32_t)r.inputs[q] << (5 * q);
			for (size_t t = 0; t < r.table.size(); ++t)
				if (r.table[t])
					R.tab[i] |= uint64_t(1) << t;
		}
		timed_launch(c, "k_flip_mask", 4.0 * n, [&] {
			k_flip_mask_small<<<nb, table_block, 0, c.stream>>>(N, n, R, g.up_ok, g.down_ok, g.mask.get(),
																block_deg.get());
		});
	}
	else
		timed_launch(c, "k_flip_mask", 4.0 * n, [&] {
			k_flip_mask<<<nb, table_block, 0, c.stream>>>(N, n, g.rules.get(), g.rule_inputs.get(),
														  g.rule_tables.get(), (int)tabs.size(), g.up_ok, g.down_ok,
														  g.mask.get(), block_deg.get());
		});
	k_scan_single<<<1, 1024, 0, c.stream>>>(block_deg.get(), nb);
	ESL_HIP(hipGetLastError());
	int32_t total = 0;
	ESL_HIP(hipMemcpyAsync(&total, block_deg.get() + nb, sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
	ESL_HIP(hipStreamSynchronize(c.stream));
	g.nnz = total;
	g.rows.ensure(total);
	g.cols.ensure(total);
	const size_t lds = sizeof(int32_t) * 2 * table_block * N;
	// algorithmic bytes: mask read (4n), indptr (4(n+1)), rows and cols (4 nnz each) written
	timed_launch(c, "k_table_fill", 4.0 * n + 4.0 * (n + 1) + 8.0 * (double)total, [&] {
		k_table_fill<<<nb, table_block, lds, c.stream>>>(N, n, g.mask.get(), block_deg.get(), g.indptr.get(),
														 g.rows.get(), g.cols.get());
	});
	ESL_HIP(hipGetLastError());
	ESL_HIP(hipEventRecord(c.ev[1], c.stream));
	ESL_HIP(hipEventSynchronize(c.ev[1]));
	float ms = 0;
	ESL_HIP(hipEventElapsedTime(&ms, c.ev[0], c.ev[1]));
	g.stage_ms[0] = ms;
}

// ---- device-to-device copy: 16-byte vectors when both ends allow it, bytes otherwise
__global__ __launch_bounds__(256) void k_copy16(int64_t count, const uint4* __restrict__ src, uint4* __restrict__ dst)
{
	for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (int64_t)gridDim.x * 256)
		dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_copy1(int64_t count, const uint8_t* __restrict__ src, uint8_t* __restrict__ dst)
{
	for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (int64_t)gridDim.x * 256)
		dst[i] = src[i];
}

void device_copy(ctx_t& c, void* dst, const void* src, size_t bytes)
{
	if (!bytes)
		return;
	const bool vec = ((uintptr_t)dst % 16 == 0) && ((uintptr_t)src % 16 == 0) && bytes % 16 == 0;
	const int64_t count = vec ? (int64_t)(bytes / 16) : (int64_t)bytes;
	const int blocks = (int)std::min<int64_t>((count + 255) / 256, 256 * 64);
	if (vec)
		k_copy16<<<blocks, 256, 0, c.stream>>>(count, (const uint4*)src, (uint4*)dst);
	else
		k_copy1<<<blocks, 256, 0, c.stream>>>(count, (const uint8_t*)src, (uint8_t*)dst);
	ESL_HIP(hipGetLastError());
}

} // namespace esl
