// Surrogate MLP: log-rates -> distribution over the terminal-SCC states.
//
// No counterpart in the reference.  A small fully connected network
// (n_in -> H -> ... -> H -> n_out, GELU between hidden layers, softmax at the
// end) trained with cross-entropy against stationary vectors the pipeline
// computes, and used to warm-start the terminal-kernel solves.
//
// Every matrix product runs on the matrix cores: operands are rounded to BF16
// and multiplied with v_mfma_f32_32x32x16_bf16, accumulating in FP32.  Master
// weights, activations and Adam moments are FP32.
//
// GEMM: a workgroup of 4 waves computes a 64 x 64 tile of C.  Each K stage
// (BK = 128, or 32 for products with K < 128) brings a 64 x BK slice of A and
// a BK x 64 slice of B from global memory into registers (16-byte loads along
// whichever index is contiguous when the operand allows, single floats
// otherwise; interior tiles without bounds tests), rounds it to BF16 once and
// stores it in LDS as rows of BK k values (A by m, B by n); the four waves then
// read their 32 x 32 sub-tiles' operands from LDS with 16-byte reads and issue
// BK / 16 MFMA 32x32x16 each.  The next stage's global loads are in flight
// while the current stage is multiplied (LDS double buffer).  Operand element
// (m, k) of A is A[m * sam + k * sak] and (k, n) of B is B[k * sbk + n * sbn],
// so the same kernel serves X W, dZ W^T and X^T dZ.  Products with few output
// tiles and a long K (dZ W^T of the widest layer) are split over gridDim.z;
// each split writes FP32 partial tiles that k_gemm_reduce sums in split order
// (deterministic) before the epilogue.  The epilogues (bias + GELU, bias,
// times GELU'(Z)) are fused, and what they read is loaded before the K loop.
// Register layout of the instruction (wave64): lane l holds
// A[l % 32][8 (l / 32) + 0..7], B[8 (l / 32) + 0..7][l % 32] and C rows
// 8 (i / 4) + 4 (l / 32) + (i % 4), column l % 32, for i = 0..15.

#include <algorithm>
#include <cmath>
#include <functional>
#include <memory>
#include <random>

#include "common.hip.hpp"
#include "esl_internal.hpp"
#include "../../include/esl_mlp.h"

namespace esl {

typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8_t;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16_t;

__device__ __forceinline__ short to_bf16(float x)
{
	// round to nearest even (finite inputs)
	uint32_t u = __float_as_uint(x);
	u += 0x7fffu + ((u >> 16) & 1u);
	return (short)(u >> 16);
}

constexpr int gemm_block = 256;
constexpr int gemm_max_splits = 8;

// LDS row of a staged operand: BK values of k in BF16 plus 8 of padding (rows
// stay 16-byte aligned and consecutive rows start in different banks)
template <int BK>
constexpr int gemm_ld = BK + 8;

enum epilogue_t
{
	EPI_STORE = 0,	   // C = A B
	EPI_BIAS_GELU = 1, // Z = A B + bias; C = gelu(Z)   (Z and C both stored)
	EPI_BIAS = 2,	   // C = A B + bias
	EPI_GELU_BWD = 3   // C = (A B) * gelu'(Z)          (Z: the layer's pre-activations, read)
};

// GELU, tanh form: 0.5 z (1 + tanh(u)), u = sqrt(2/pi) (z + 0.044715 z^3),
// evaluated as z sigma(2u) with sigma(x) = 1 / (1 + exp(-x)) -- one hardware
// exponential and one hardware reciprocal (about 1 ulp each) instead of a tanh
// call, which dominated the epilogue of the short products;
// 1 + tanh(u) = 2 sigma(2u) and 1 - tanh(u)^2 = 4 sigma(2u) (1 - sigma(2u))
// give the derivative.
__device__ __forceinline__ float gelu_sigma(float z, float& du)
{
	const float c = 0.7978845608028654f; // sqrt(2/pi)
	const float u = c * (z + 0.044715f * z * z * z);
	du = c * (1.0f + 3.0f * 0.044715f * z * z);
	return __builtin_amdgcn_rcpf(1.0f + __expf(-2.0f * u));
}

__device__ __forceinline__ float gelu_f(float z)
{
	float du;
	return z * gelu_sigma(z, du);
}

__device__ __forceinline__ float gelu_grad_f(float z)
{
	float du;
	const float sg = gelu_sigma(z, du);
	return sg + 2.0f * z * sg * (1.0f - sg) * du;
}

// bv: bias[cc] (EPI_BIAS_GELU, EPI_BIAS); zv: Z[r][cc] (EPI_GELU_BWD)
__device__ __forceinline__ void gemm_epilogue(int epi, int N, int r, int cc, float v, float bv, float zv,
											  float* __restrict__ C, float* __restrict__ Z)
{
	const int64_t o = (int64_t)r * N + cc;
	if (epi == EPI_BIAS_GELU)
	{
		v += bv;
		Z[o] = v;
		C[o] = gelu_f(v);
	}
	else if (epi == EPI_BIAS)
		C[o] = v + bv;
	else if (epi == EPI_GELU_BWD)
		C[o] = v * gelu_grad_f(zv);
	else
		C[o] = v;
}

// One K stage of an operand, 64 rows x BK k, whose element (r, k) is
// P[r * sr + k * sk], held in registers between its global loads and its BF16
// store to LDS (BK / 4 FP32 values per thread).  KC: k is the contiguous index
// (sk == 1); otherwise r is (sr == 1).
//
// VEC (16-byte aligned operand, contiguous extent and the other stride
// multiples of 4): 16-byte loads.  KC: BK / 4 lanes take consecutive k quads
// of one row and store them as 8 bytes of BF16.  Otherwise a lane takes a quad
// of consecutive rows at k and at k + 1 and stores the four (k, k + 1) pairs as
// packed BF16 words.  Without VEC the loads are single floats (KC: BK / 2
// lanes take consecutive k pairs of a row; otherwise the 64 lanes of a wave
// take consecutive rows of one k pair).  Every load instruction of a wave reads
// whole cache lines either way.  CHECK (edge tiles only): values outside
// [0, R) x [k, kend) read as zero; interior tiles load without tests.
template <bool KC, bool VEC, int BK>
struct stage_t
{
	static constexpr int nf = BK / 4; // FP32 values per thread
	float v[nf];

	static __device__ __forceinline__ int t() { return (int)threadIdx.x; }

	template <bool CHECK>
	__device__ __forceinline__ void load(const float* __restrict__ P, int64_t sr, int64_t sk, int r0, int R, int k0,
										 int kend)
	{
		if constexpr (VEC && KC)
		{
#pragma unroll
			for (int q = 0; q < nf / 4; ++q)
			{
				const int r = r0 + t() / (BK / 4) + (1024 / BK) * q, k = k0 + 4 * (t() % (BK / 4));
				float4 x = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
				if (!CHECK || (r < R && k < kend))
					x = *reinterpret_cast<const float4*>(P + (int64_t)r * sr + k);
				v[4 * q] = x.x, v[4 * q + 1] = x.y, v[4 * q + 2] = x.z, v[4 * q + 3] = x.w;
			}
		}
		else if constexpr (VEC)
		{
#pragma unroll
			for (int q = 0; q < nf / 8; ++q)
			{
				const int r = r0 + 4 * (t() & 15), k = k0 + 2 * (t() / 16 + 16 * q);
				const float* p = P + r + (int64_t)k * sk;
				float4 x = make_float4(0.0f, 0.0f, 0.0f, 0.0f), y = x;
				if (!CHECK || (r < R && k < kend))
					x = *reinterpret_cast<const float4*>(p);
				if (!CHECK || (r < R && k + 1 < kend))
					y = *reinterpret_cast<const float4*>(p + sk);
				v[8 * q] = x.x, v[8 * q + 1] = y.x, v[8 * q + 2] = x.y, v[8 * q + 3] = y.y;
				v[8 * q + 4] = x.z, v[8 * q + 5] = y.z, v[8 * q + 6] = x.w, v[8 * q + 7] = y.w;
			}
		}
		else
		{
#pragma unroll
			for (int q = 0; q < nf / 2; ++q)
			{
				const int r = r0 + (KC ? t() / (BK / 2) + (512 / BK) * q : (t() & 63));
				const int k = k0 + (KC ? 2 * (t() % (BK / 2)) : 2 * ((t() >> 6) + 4 * q));
				const float* p = P + (int64_t)r * sr + (int64_t)k * sk;
				v[2 * q] = (!CHECK || (r < R && k < kend)) ? p[0] : 0.0f;
				v[2 * q + 1] = (!CHECK || (r < R && k + 1 < kend)) ? p[sk] : 0.0f;
			}
		}
	}

	__device__ __forceinline__ void store(short* __restrict__ s) const
	{
		constexpr int ld = gemm_ld<BK>;
		if constexpr (VEC && KC)
		{
#pragma unroll
			for (int q = 0; q < nf / 4; ++q)
			{
				const int row = t() / (BK / 4) + (1024 / BK) * q, k = 4 * (t() % (BK / 4));
				uint2 w;
				w.x = (uint32_t)(uint16_t)to_bf16(v[4 * q]) | ((uint32_t)(uint16_t)to_bf16(v[4 * q + 1]) << 16);
				w.y = (uint32_t)(uint16_t)to_bf16(v[4 * q + 2]) | ((uint32_t)(uint16_t)to_bf16(v[4 * q + 3]) << 16);
				*reinterpret_cast<uint2*>(s + row * ld + k) = w;
			}
		}
		else if constexpr (VEC)
		{
#pragma unroll
			for (int q = 0; q < nf / 8; ++q)
			{
				const int row = 4 * (t() & 15), k = 2 * (t() / 16 + 16 * q);
#pragma unroll
				for (int i = 0; i < 4; ++i)
					*reinterpret_cast<uint32_t*>(s + (row + i) * ld + k) =
						(uint32_t)(uint16_t)to_bf16(v[8 * q + 2 * i]) |
						((uint32_t)(uint16_t)to_bf16(v[8 * q + 2 * i + 1]) << 16);
			}
		}
		else
		{
#pragma unroll
			for (int q = 0; q < nf / 2; ++q)
			{
				const int row = KC ? t() / (BK / 2) + (512 / BK) * q : (t() & 63);
				const int k = KC ? 2 * (t() % (BK / 2)) : 2 * ((t() >> 6) + 4 * q);
				*reinterpret_cast<uint32_t*>(s + row * ld + k) =
					(uint32_t)(uint16_t)to_bf16(v[2 * q]) | ((uint32_t)(uint16_t)to_bf16(v[2 * q + 1]) << 16);
			}
		}
	}
};

// C (M x N) = A (M x K) B (K x N) over the K range [z k_split, (z + 1) k_split)
// of split z = blockIdx.z; with part set the raw FP32 tile goes to
// part[z][M][N] for k_gemm_reduce, else through the epilogue.  AKC / BKC: the
// operand's k index is the contiguous one (sak == 1 / sbk == 1); VEC: both
// operands take 16-byte loads.
template <bool AKC, bool BKC, bool VEC, int BK>
__global__ __launch_bounds__(gemm_block) void k_gemm_bf16(int M, int N, int K, int k_split,
														  const float* __restrict__ A, int64_t sam, int64_t sak,
														  const float* __restrict__ Bm, int64_t sbk, int64_t sbn,
														  const float* __restrict__ bias, float* __restrict__ C,
														  float* __restrict__ Z, int epi, float* __restrict__ part)
{
	constexpr int ld = gemm_ld<BK>;
	__shared__ __align__(16) short sA[2][64 * ld];
	__shared__ __align__(16) short sB[2][64 * ld];
	const int w = threadIdx.x / wave;
	const int l = lane_id();
	const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
	const int kb = blockIdx.z * k_split, ke = min(K, kb + k_split);
	const bool rows_full = m0 + 64 <= M && n0 + 64 <= N;
	stage_t<AKC, VEC, BK> ga;
	stage_t<BKC, VEC, BK> gb; // B's rows are its columns n: element (n, k) = Bm[n * sbn + k * sbk]
	auto fetch = [&](int k0) {
		if (rows_full && k0 + BK <= ke)
		{
			ga.template load<false>(A, sam, sak, m0, M, k0, ke);
			gb.template load<false>(Bm, sbn, sbk, n0, N, k0, ke);
		}
		else
		{
			ga.template load<true>(A, sam, sak, m0, M, k0, ke);
			gb.template load<true>(Bm, sbn, sbk, n0, N, k0, ke);
		}
	};
	// what the epilogue reads is loaded first, so its latency overlaps the products
	const int mw = m0 + (w >> 1) * 32, nw = n0 + (w & 1) * 32;
	const int cc = nw + (l & 31); // this lane's column of C
	float bv = 0.0f, zv[16];
	if (!part && (epi == EPI_BIAS_GELU || epi == EPI_BIAS) && cc < N)
		bv = bias[cc];
#pragma unroll
	for (int i = 0; i < 16; ++i)
	{
		const int r = mw + (i >> 2) * 8 + (l >> 5) * 4 + (i & 3);
		zv[i] = (!part && epi == EPI_GELU_BWD && r < M && cc < N) ? Z[(int64_t)r * N + cc] : 0.0f;
	}
	fetch(kb);
	ga.store(sA[0]);
	gb.store(sB[0]);
	__syncthreads();
	// this lane's operand rows in the staged tiles: A row m, B row n, 8 k from 8 (l / 32)
	const int ao = ((w >> 1) * 32 + (l & 31)) * ld + 8 * (l >> 5);
	const int bo = ((w & 1) * 32 + (l & 31)) * ld + 8 * (l >> 5);
	f32x16_t acc;
	for (int i = 0; i < 16; ++i)
		acc[i] = 0.0f;
	int buf = 0;
	for (int k0 = kb; k0 < ke; k0 += BK)
	{
		const bool more = k0 + BK < ke;
		if (more)
			fetch(k0 + BK);
		const int kn = min(BK, ke - k0); // MFMA steps past the range would multiply zeros
#pragma unroll
		for (int ks = 0; ks < BK; ks += 16)
		{
			if (ks >= kn)
				break;
			const bf16x8_t a = *reinterpret_cast<const bf16x8_t*>(sA[buf] + ao + ks);
			const bf16x8_t b = *reinterpret_cast<const bf16x8_t*>(sB[buf] + bo + ks);
			acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
		}
		if (more)
		{
			// buffer buf ^ 1 was last read before the previous barrier
			ga.store(sA[buf ^ 1]);
			gb.store(sB[buf ^ 1]);
		}
		__syncthreads();
		buf ^= 1;
	}
#pragma unroll
	for (int i = 0; i < 16; ++i)
	{
		const int r = mw + (i >> 2) * 8 + (l >> 5) * 4 + (i & 3);
		if (r >= M || cc >= N)
			continue;
		if (part)
			part[((int64_t)blockIdx.z * M + r) * N + cc] = acc[i];
		else
			gemm_epilogue(epi, N, r, cc, acc[i], bv, zv[i], C, Z);
	}
}

// sums the splits' partial tiles in split order, then the epilogue
__global__ __launch_bounds__(256) void k_gemm_reduce(int M, int N, int splits, const float* __restrict__ part,
													 const float* __restrict__ bias, float* __restrict__ C,
													 float* __restrict__ Z, int epi)
{
	const int64_t MN = (int64_t)M * N;
	const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if (i >= MN)
		return;
	float v = 0.0f;
	for (int s = 0; s < splits; ++s)
		v += part[s * MN + i];
	const int cc = (int)(i % N);
	const float bv = (epi == EPI_BIAS_GELU || epi == EPI_BIAS) ? bias[cc] : 0.0f;
	const float zv = epi == EPI_GELU_BWD ? Z[i] : 0.0f;
	gemm_epilogue(epi, N, (int)(i / N), cc, v, bv, zv, C, Z);
}

// b = (double) a
__global__ __launch_bounds__(256) void k_widen(int64_t count, const float* __restrict__ a, double* __restrict__ b)
{
	const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if (i < count)
		b[i] = (double)a[i];
}

// per row: P = softmax(Z); loss += -sum t log p; dZ = (P - T) / rows.  With
// NV > 0 (n <= 256 NV) the row's logits and targets are read once into
// registers, NV per thread; NV = 0 reads them from memory in each pass.
template <int NV>
__global__ __launch_bounds__(256) void k_softmax_xent(int rows, int n, const float* __restrict__ Z,
													  const float* __restrict__ Tg, float* __restrict__ P,
													  float* __restrict__ dZ, float* __restrict__ loss_rows)
{
	__shared__ float s_red[256 / wave + 1];
	const int r = blockIdx.x;
	const float* z = Z + (int64_t)r * n;
	const float* tg = Tg ? Tg + (int64_t)r * n : nullptr;
	constexpr int nc = NV > 0 ? NV : 1;
	float zc[nc], tc[nc];
	float mx = -INFINITY;
	if constexpr (NV > 0)
	{
#pragma unroll
		for (int j = 0; j < NV; ++j)
		{
			const int i = threadIdx.x + 256 * j;
			zc[j] = i < n ? z[i] : -INFINITY;
			tc[j] = (tg && i < n) ? tg[i] : 0.0f;
			mx = fmaxf(mx, zc[j]);
		}
	}
	else
		for (int i = threadIdx.x; i < n; i += 256)
			mx = fmaxf(mx, z[i]);
	mx = wave_max(mx);
	if (lane_id() == 0)
		s_red[threadIdx.x / wave] = mx;
	__syncthreads();
	mx = s_red[0];
	for (int i = 1; i < 256 / wave; ++i)
		mx = fmaxf(mx, s_red[i]);
	__syncthreads();
	float s = 0.0f;
	if constexpr (NV > 0)
	{
#pragma unroll
		for (int j = 0; j < NV; ++j)
			if (threadIdx.x + 256 * j < n)
				s += __expf(zc[j] - mx);
	}
	else
		for (int i = threadIdx.x; i < n; i += 256)
			s += __expf(z[i] - mx);
	s = block_sum<256>(s, s_red);
	const float lse = mx + logf(s);
	float l = 0.0f;
	auto out = [&](int i, float zi, float ti) {
		const float p = __expf(zi - lse);
		P[(int64_t)r * n + i] = p;
		if (tg)
		{
			if (dZ)
				dZ[(int64_t)r * n + i] = (p - ti) / (float)rows;
			if (ti > 0.0f)
				l -= ti * (zi - lse);
		}
	};
	if constexpr (NV > 0)
	{
#pragma unroll
		for (int j = 0; j < NV; ++j)
			if (threadIdx.x + 256 * j < n)
				out(threadIdx.x + 256 * j, zc[j], tc[j]);
	}
	else
		for (int i = threadIdx.x; i < n; i += 256)
			out(i, z[i], tg ? tg[i] : 0.0f);
	l = block_sum<256>(l, s_red);
	if (threadIdx.x == 0 && loss_rows)
		loss_rows[r] = l;
}

// db[c] = sum over rows of dZ[r][c]: 32 columns per workgroup, 8 slices of
// consecutive rows summed by the 8 half-waves, then the slices in order
// (a fixed order: the sums are reproducible)
constexpr int colsum_cols = 32;

__global__ __launch_bounds__(256) void k_colsum(int rows, int n, const float* __restrict__ dZ, float* __restrict__ db)
{
	__shared__ float s_part[8][colsum_cols];
	const int cl = threadIdx.x % colsum_cols, sl = threadIdx.x / colsum_cols;
	const int c = blockIdx.x * colsum_cols + cl;
	const int per = (rows + 7) / 8;
	const int r0 = sl * per, r1 = min(rows, r0 + per);
	float acc = 0.0f;
	if (c < n)
	{
#pragma unroll 8
		for (int r = r0; r < r1; ++r)
			acc += dZ[(int64_t)r * n + c];
	}
	s_part[sl][cl] = acc;
	__syncthreads();
	if (sl == 0 && c < n)
	{
		float t = 0.0f;
		for (int i = 0; i < 8; ++i)
			t += s_part[i][cl];
		db[c] = t;
	}
}

__global__ __launch_bounds__(256) void k_adam(int64_t count, float* __restrict__ p, const float* __restrict__ g,
											  float* __restrict__ m, float* __restrict__ v, float lr, float b1,
											  float b2, float eps, float c1, float c2)
{
	const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if (i >= count)
		return;
	const float gi = g[i];
	const float mi = b1 * m[i] + (1.0f - b1) * gi;
	const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
	m[i] = mi;
	v[i] = vi;
	p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// out[0] = sum of x (add: out[0] += sum of x)
__global__ __launch_bounds__(256) void k_sum_f(int n, const float* __restrict__ x, float* __restrict__ out, int add)
{
	__shared__ float s_red[256 / wave + 1];
	float s = 0.0f;
	for (int i = threadIdx.x; i < n; i += 256)
		s += x[i];
	s = block_sum<256>(s, s_red);
	if (threadIdx.x == 0)
		out[0] = add ? out[0] + s : s;
}

} // namespace esl

// ============================================================================
// host side

struct esl_mlp
{
	int n_in = 0, hidden = 0, layers = 0, n_out = 0;
	std::vector<int> dims;			 // n_in, H, ..., H, n_out
	std::vector<int64_t> w_off, b_off; // offsets into the parameter buffer
	int64_t n_params = 0;
	esl::dbuf<float> params, grads, m1, m2;
	int64_t step = 0;
	int cap_rows = 0;
	esl::dbuf<float> act, pre, dact, probs, dlog, loss_rows, loss;
	esl::dbuf<float> part;		  // split-K partial tiles of the GEMMs
	std::vector<int64_t> act_off; // activations of each layer (input copy at 0)
};

namespace {

esl_status mlp_guard(esl_ctx* ctx, const std::function<void()>& f)
{
	// same contract as the other entry points: no exception crosses the ABI
	try
	{
		f();
		return ESL_OK;
	}
	catch (const esl::error& e)
	{
		if (ctx)
			esl::ctx_impl(ctx).last_error = e.what();
		return (esl_status)e.code;
	}
	catch (const std::exception& e)
	{
		if (ctx)
			esl::ctx_impl(ctx).last_error = e.what();
		return ESL_E_ARG;
	}
}

void ensure_rows(esl_mlp* m, int rows)
{
	if (rows <= m->cap_rows)
		return;
	int64_t tot = 0;
	m->act_off.assign(m->dims.size(), 0);
	for (size_t i = 0; i < m->dims.size(); ++i)
	{
		m->act_off[i] = tot;
		tot += (int64_t)rows * m->dims[i];
	}
	m->act.ensure(tot);
	m->pre.ensure(tot);
	m->dact.ensure(tot);
	m->probs.ensure((size_t)rows * m->n_out);
	m->dlog.ensure((size_t)rows * m->n_out);
	m->loss_rows.ensure(rows);
	m->loss.ensure(1);
	m->cap_rows = rows;
}

template <bool VEC, int BK>
void launch_gemm(esl::ctx_t& c, dim3 grid, int M, int N, int K, int k_split, const float* A, int64_t sam,
				 int64_t sak, const float* B, int64_t sbk, int64_t sbn, const float* bias, float* C, float* Z, int epi,
				 float* part)
{
#define ESL_GEMM(AKC, BKC)                                                                                             \
	esl::k_gemm_bf16<AKC, BKC, VEC, BK><<<grid, esl::gemm_block, 0, c.stream>>>(M, N, K, k_split, A, sam, sak, B, sbk, \
																				sbn, bias, C, Z, epi, part)
	if (sak == 1 && sbk == 1)
		ESL_GEMM(true, true);
	else if (sak == 1)
		ESL_GEMM(true, false);
	else if (sbk == 1)
		ESL_GEMM(false, true);
	else
		ESL_GEMM(false, false);
#undef ESL_GEMM
}

// an operand of R rows x K (element (r, k) at P[r sr + k sk]) can be read in
// aligned quads along its contiguous index
bool quad_loads(const float* P, int R, int K, int64_t sr, int64_t sk)
{
	if ((uintptr_t)P % 16)
		return false;
	if (sk == 1)
		return sr % 4 == 0 && K % 4 == 0;
	return sr == 1 && sk % 4 == 0 && R % 4 == 0;
}

// C = A B with the given epilogue.  Products with few 64 x 64 output tiles and
// K >= 512 are split along K so that the launch has up to 256 workgroups of at
// least 256 k each.  K stages of 128 (32 for short products) keep enough loads
// in flight per stage that the few stages of these products are not each a
// full memory latency.
void gemm(esl::ctx_t& c, esl_mlp* m, int M, int N, int K, const float* A, int64_t sam, int64_t sak, const float* B,
		  int64_t sbk, int64_t sbn, const float* bias, float* C, float* Z, int epi)
{
	const int tiles = esl::ceil_div(M, 64) * esl::ceil_div(N, 64);
	int splits = std::max(1, std::min({ K / 256, 256 / tiles, esl::gemm_max_splits }));
	const int kper = esl::ceil_div(K, splits);
	const int bk = kper >= 128 ? 128 : 32;
	const int k_split = esl::ceil_div(esl::ceil_div(K, splits), bk) * bk;
	splits = esl::ceil_div(K, k_split);
	float* part = nullptr;
	if (splits > 1)
	{
		m->part.ensure((size_t)splits * M * N);
		part = m->part.get();
	}
	const dim3 grid(esl::ceil_div(N, 64), esl::ceil_div(M, 64), splits);
	const double flops = 2.0 * M * N * K;
	esl::timed_launch(c, "k_gemm_bf16", flops, [&] {
		const bool vec = quad_loads(A, M, K, sam, sak) && quad_loads(B, N, K, sbn, sbk);
#define ESL_GEMM_BK(V, BK) launch_gemm<V, BK>(c, grid, M, N, K, k_split, A, sam, sak, B, sbk, sbn, bias, C, Z, epi, part)
		if (vec)
			bk == 128 ? ESL_GEMM_BK(true, 128) : ESL_GEMM_BK(true, 32);
		else
			bk == 128 ? ESL_GEMM_BK(false, 128) : ESL_GEMM_BK(false, 32);
#undef ESL_GEMM_BK
	});
	if (splits > 1)
	{
		const int64_t mn = (int64_t)M * N;
		esl::timed_launch(c, "k_gemm_reduce", 4.0 * (splits + 1) * mn, [&] {
			esl::k_gemm_reduce<<<esl::ceil_div(mn, 256), 256, 0, c.stream>>>(M, N, splits, part, bias, C, Z, epi);
		});
	}
}

void softmax_xent(esl::ctx_t& c, int rows, int n, const float* Zl, const float* T, float* P, float* dZ, float* loss_rows)
{
	if (n <= 8 * 256)
		esl::k_softmax_xent<8><<<rows, 256, 0, c.stream>>>(rows, n, Zl, T, P, dZ, loss_rows);
	else if (n <= 16 * 256)
		esl::k_softmax_xent<16><<<rows, 256, 0, c.stream>>>(rows, n, Zl, T, P, dZ, loss_rows);
	else
		esl::k_softmax_xent<0><<<rows, 256, 0, c.stream>>>(rows, n, Zl, T, P, dZ, loss_rows);
}

// forward pass on rows x n_in inputs x (read in place); logits in pre[last]
void forward(esl::ctx_t& c, esl_mlp* m, int rows, const float* x)
{
	const int L = (int)m->dims.size() - 1;
	for (int l = 0; l < L; ++l)
	{
		const int K = m->dims[l], N = m->dims[l + 1];
		const float* X = l == 0 ? x : m->act.get() + m->act_off[l];
		const float* W = m->params.get() + m->w_off[l];
		const float* b = m->params.get() + m->b_off[l];
		float* Y = m->act.get() + m->act_off[l + 1];
		float* Z = m->pre.get() + m->act_off[l + 1];
		if (l < L - 1)
			gemm(c, m, rows, N, K, X, K, 1, W, N, 1, b, Y, Z, esl::EPI_BIAS_GELU);
		else
			gemm(c, m, rows, N, K, X, K, 1, W, N, 1, b, Z, nullptr, esl::EPI_BIAS);
	}
}

} // namespace

extern "C" {

esl_status esl_mlp_create(esl_ctx* ctx, int n_in, int hidden, int hidden_layers, int n_out, uint64_t seed,
						  esl_mlp** out)
{
	return mlp_guard(ctx, [&] {
		if (!ctx || !out || n_in < 1 || hidden < 1 || hidden_layers < 1 || n_out < 2)
			throw esl::error(ESL_E_ARG, "bad MLP shape");
		ESL_HIP(hipSetDevice(esl::ctx_impl(ctx).device));
		auto m = std::make_unique<esl_mlp>();
		m->n_in = n_in;
		m->hidden = hidden;
		m->layers = hidden_layers;
		m->n_out = n_out;
		m->dims.push_back(n_in);
		for (int i = 0; i < hidden_layers; ++i)
			m->dims.push_back(hidden);
		m->dims.push_back(n_out);
		int64_t off = 0;
		for (size_t l = 0; l + 1 < m->dims.size(); ++l)
		{
			m->w_off.push_back(off);
			off += (int64_t)m->dims[l] * m->dims[l + 1];
			m->b_off.push_back(off);
			off += m->dims[l + 1];
		}
		m->n_params = off;
		m->params.ensure(off);
		m->grads.ensure(off);
		m->m1.ensure(off);
		m->m2.ensure(off);
		// He-style init from a seeded std::mt19937_64, zero biases
		std::vector<float> h(off, 0.0f);
		std::mt19937_64 rng(seed);
		for (size_t l = 0; l + 1 < m->dims.size(); ++l)
		{
			std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / m->dims[l]));
			for (int64_t i = 0; i < (int64_t)m->dims[l] * m->dims[l + 1]; ++i)
				h[m->w_off[l] + i] = (float)nd(rng);
		}
		ESL_HIP(hipMemcpy(m->params.get(), h.data(), off * sizeof(float), hipMemcpyHostToDevice));
		ESL_HIP(hipMemset(m->m1.get(), 0, off * sizeof(float)));
		ESL_HIP(hipMemset(m->m2.get(), 0, off * sizeof(float)));
		*out = m.release();
	});
}

void esl_mlp_destroy(esl_mlp* m) { delete m; }

int64_t esl_mlp_num_params(const esl_mlp* m) { return m ? m->n_params : 0; }

esl_status esl_mlp_get_params(esl_ctx* ctx, const esl_mlp* m, float* host_out)
{
	return mlp_guard(ctx, [&] {
		if (!ctx || !m || !host_out)
			throw esl::error(ESL_E_ARG, "null argument");
		ESL_HIP(hipMemcpy(host_out, m->params.get(), m->n_params * sizeof(float), hipMemcpyDeviceToHost));
	});
}

esl_status esl_mlp_set_params(esl_ctx* ctx, esl_mlp* m, const float* host_in)
{
	return mlp_guard(ctx, [&] {
		if (!ctx || !m || !host_in)
			throw esl::error(ESL_E_ARG, "null argument");
		ESL_HIP(hipMemcpy(m->params.get(), host_in, m->n_params * sizeof(float), hipMemcpyHostToDevice));
	});
}

esl_status esl_mlp_forward(esl_ctx* ctx, esl_mlp* m, int rows, const float* d_x, float* d_probs)
{
	return mlp_guard(ctx, [&] {
		if (!ctx || !m || !d_x || !d_probs || rows < 1)
			throw esl::error(ESL_E_ARG, "null argument");
		esl::ctx_t& c = esl::ctx_impl(ctx);
		ensure_rows(m, rows);
		forward(c, m, rows, d_x);
		const int L = (int)m->dims.size() - 1;
		softmax_xent(c, rows, m->n_out, m->pre.get() + m->act_off[L], nullptr, d_probs, nullptr, nullptr);
		ESL_HIP(hipGetLastError());
		ESL_HIP(hipStreamSynchronize(c.stream));
	});
}

esl_status esl_mlp_predict_warm(esl_ctx* ctx, esl_mlp* m, int rows, const float* d_x, double* d_warm)
{
	return mlp_guard(ctx, [&] {
		if (!ctx || !m || !d_x || !d_warm || rows < 1)
			throw esl::error(ESL_E_ARG, "null argument");
		esl::ctx_t& c = esl::ctx_impl(ctx);
		ensure_rows(m, rows);
		forward(c, m, rows, d_x);
		const int L = (int)m->dims.size() - 1;
		softmax_xent(c, rows, m->n_out, m->pre.get() + m->act_off[L], nullptr, m->probs.get(), nullptr, nullptr);
		const int64_t cnt = (int64_t)rows * m->n_out;
		esl::k_widen<<<esl::ceil_div(cnt, 256), 256, 0, c.stream>>>(cnt, m->probs.get(), d_warm);
		ESL_HIP(hipGetLastError());
		ESL_HIP(hipStreamSynchronize(c.stream));
	});
}

// gradients of the mean cross-entropy into m->grads; the summed loss of the
// rows goes to m->loss[0] (added to it with accumulate).  Enqueued only.
static void mlp_enqueue_backward(esl::ctx_t& c, esl_mlp* m, int rows, const float* d_x, const float* d_t,
								 bool accumulate)
{
	ensure_rows(m, rows);
	forward(c, m, rows, d_x);
	const int L = (int)m->dims.size() - 1;
	softmax_xent(c, rows, m->n_out, m->pre.get() + m->act_off[L], d_t, m->probs.get(), m->dact.get() + m->act_off[L],
				 m->loss_rows.get());
	for (int l = L - 1; l >= 0; --l)
	{
		const int K = m->dims[l], N = m->dims[l + 1];
		float* dZ = m->dact.get() + m->act_off[l + 1]; // rows x N (already multiplied by gelu')
		const float* X = l == 0 ? d_x : m->act.get() + m->act_off[l];
		float* dW = m->grads.get() + m->w_off[l];
		float* db = m->grads.get() + m->b_off[l];
		// dW = X^T dZ  (K x N)
		gemm(c, m, K, N, rows, X, 1, K, dZ, N, 1, nullptr, dW, nullptr, esl::EPI_STORE);
		esl::k_colsum<<<esl::ceil_div(N, esl::colsum_cols), 256, 0, c.stream>>>(rows, N, dZ, db);
		if (l > 0)
		{
			// dX = (dZ W^T) * gelu'(Z_l)  (rows x K)
			const float* W = m->params.get() + m->w_off[l];
			float* dX = m->dact.get() + m->act_off[l];
			gemm(c, m, rows, K, N, dZ, N, 1, W, 1, N, nullptr, dX, m->pre.get() + m->act_off[l], esl::EPI_GELU_BWD);
		}
	}
	esl::k_sum_f<<<1, 256, 0, c.stream>>>(rows, m->loss_rows.get(), m->loss.get(), accumulate ? 1 : 0);
	ESL_HIP(hipGetLastError());
}

// m->loss[0] to the host (waits for the stream)
static float mlp_read_loss(esl::ctx_t& c, esl_mlp* m)
{
	float loss = 0.0f;
	ESL_HIP(hipMemcpyAsync(&loss, m->loss.get(), sizeof(float), hipMemcpyDeviceToHost, c.stream));
	ESL_HIP(hipStreamSynchronize(c.stream));
	return loss;
}

// one Adam step on m->grads (beta1 0.9, beta2 0.999, eps 1e-8), enqueued
static void mlp_enqueue_adam(esl::ctx_t& c, esl_mlp* m, float lr)
{
	m->step++;
	const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
	const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
	esl::k_adam<<<esl::ceil_div(m->n_params, 256), 256, 0, c.stream>>>(m->n_params, m->params.get(), m->grads.get(),
																		m->m1.get(), m->m2.get(), lr, b1, b2, eps,
																		c1, c2);
	ESL_HIP(hipGetLastError());
}

esl_status esl_mlp_gradients(esl_ctx* ctx, esl_mlp* m, int rows, const float* d_x, const float* d_target,
							 float* host_grads, float* loss_out)
{
	return mlp_guard(ctx, [&] {
		if (!ctx || !m || !d_x || !d_target || rows < 1)
			throw esl::error(ESL_E_ARG, "null argument");
		esl::ctx_t& c = esl::ctx_impl(ctx);
		mlp_enqueue_backward(c, m, rows, d_x, d_target, false);
		const float loss = mlp_read_loss(c, m) / rows;
		if (loss_out)
			*loss_out = loss;
		if (host_grads)
			ESL_HIP(hipMemcpy(host_grads, m->grads.get(), m->n_params * sizeof(float), hipMemcpyDeviceToHost));
	});
}

esl_status esl_mlp_train_step(esl_ctx* ctx, esl_mlp* m, int rows, const float* d_x, const float* d_target, float lr,
							  float* loss_out)
{
	return mlp_guard(ctx, [&] {
		if (!ctx || !m || !d_x || !d_target || rows < 1)
			throw esl::error(ESL_E_ARG, "null argument");
		esl::ctx_t& c = esl::ctx_impl(ctx);
		mlp_enqueue_backward(c, m, rows, d_x, d_target, false);
		mlp_enqueue_adam(c, m, lr);
		const float loss = mlp_read_loss(c, m) / rows;
		if (loss_out)
			*loss_out = loss;
	});
}

esl_status esl_mlp_train_steps(esl_ctx* ctx, esl_mlp* m, int steps, int rows, const float* d_x,
							   const float* d_target, float lr_first, float lr_last, float* mean_loss_out)
{
	return mlp_guard(ctx, [&] {
		if (!ctx || !m || !d_x || !d_target || rows < 1 || steps < 1 || !(lr_first >= 0.0f) || !(lr_last >= 0.0f))
			throw esl::error(ESL_E_ARG, "bad training-steps argument");
		esl::ctx_t& c = esl::ctx_impl(ctx);
		const double ratio = lr_first > 0.0f ? (double)lr_last / lr_first : 0.0;
		for (int j = 0; j < steps; ++j)
		{
			const float lr =
				steps == 1 || lr_first == 0.0f ? lr_first : (float)(lr_first * std::pow(ratio, (double)j / (steps - 1)));
			mlp_enqueue_backward(c, m, rows, d_x + (int64_t)j * rows * m->n_in,
								 d_target + (int64_t)j * rows * m->n_out, j > 0);
			mlp_enqueue_adam(c, m, lr);
		}
		const float loss = mlp_read_loss(c, m) / ((float)rows * steps);
		if (mean_loss_out)
			*mean_loss_out = loss;
	});
}

} // extern "C"
