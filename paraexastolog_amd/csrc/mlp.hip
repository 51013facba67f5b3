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
// GEMM: one wavefront computes a 32 x 32 tile of C with MFMA 32x32x16 (K step
// 16), a workgroup of 4 waves a 64 x 64 tile.  Operand element (m, k) of A is
// A[m * sam + k * sak] and (k, n) of B is B[k * sbk + n * sbn], so the same
// kernel serves X W, dZ W^T and X^T dZ.  Register layout of the instruction
// (wave64): lane l holds A[l % 32][8 (l / 32) + 0..7], B[8 (l / 32) + 0..7][l % 32]
// and C rows 8 (i / 4) + 4 (l / 32) + (i % 4), column l % 32, for i = 0..15.

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

enum epilogue_t
{
	EPI_STORE = 0,	   // C = A B
	EPI_BIAS_GELU = 1, // Z = A B + bias; C = gelu(Z)   (Z and C both stored)
	EPI_BIAS = 2	   // C = A B + bias
};

__device__ __forceinline__ float gelu_f(float z)
{
	const float c = 0.7978845608028654f; // sqrt(2/pi)
	return 0.5f * z * (1.0f + tanhf(c * (z + 0.044715f * z * z * z)));
}

__device__ __forceinline__ float gelu_grad_f(float z)
{
	const float c = 0.7978845608028654f;
	const float u = c * (z + 0.044715f * z * z * z);
	const float t = tanhf(u);
	const float du = c * (1.0f + 3.0f * 0.044715f * z * z);
	return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * du;
}

__global__ __launch_bounds__(gemm_block) void k_gemm_bf16(int M, int N, int K, const float* __restrict__ A,
														  int64_t sam, int64_t sak, const float* __restrict__ Bm,
														  int64_t sbk, int64_t sbn, const float* __restrict__ bias,
														  float* __restrict__ C, float* __restrict__ Z, int epi)
{
	const int w = threadIdx.x / wave;
	const int l = lane_id();
	const int m0 = blockIdx.y * 64 + (w >> 1) * 32;
	const int n0 = blockIdx.x * 64 + (w & 1) * 32;
	const int row = m0 + (l & 31);
	const int col = n0 + (l & 31);
	const int kh = (l >> 5) * 8;
	f32x16_t acc;
	for (int i = 0; i < 16; ++i)
		acc[i] = 0.0f;
	for (int k0 = 0; k0 < K; k0 += 16)
	{
		bf16x8_t a, b;
		for (int q = 0; q < 8; ++q)
		{
			const int k = k0 + kh + q;
			const float av = (row < M && k < K) ? A[(int64_t)row * sam + (int64_t)k * sak] : 0.0f;
			const float bv = (col < N && k < K) ? Bm[(int64_t)k * sbk + (int64_t)col * sbn] : 0.0f;
			a[q] = to_bf16(av);
			b[q] = to_bf16(bv);
		}
		acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
	}
	for (int i = 0; i < 16; ++i)
	{
		const int r = m0 + (i >> 2) * 8 + (l >> 5) * 4 + (i & 3);
		const int cc = n0 + (l & 31);
		if (r >= M || cc >= N)
			continue;
		float v = acc[i];
		if (epi == EPI_BIAS_GELU)
		{
			v += bias[cc];
			Z[(int64_t)r * N + cc] = v;
			C[(int64_t)r * N + cc] = gelu_f(v);
		}
		else if (epi == EPI_BIAS)
			C[(int64_t)r * N + cc] = v + bias[cc];
		else
			C[(int64_t)r * N + cc] = v;
	}
}

// per row: P = softmax(Z); loss += -sum t log p; dZ = (P - T) / rows
__global__ __launch_bounds__(256) void k_widen(int64_t count, const float* __restrict__ a, double* __restrict__ b)
{
	const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if (i < count)
		b[i] = (double)a[i];
}

__global__ __launch_bounds__(256) void k_softmax_xent(int rows, int n, const float* __restrict__ Z,
													  const float* __restrict__ Tg, float* __restrict__ P,
													  float* __restrict__ dZ, float* __restrict__ loss_rows)
{
	__shared__ float s_red[256 / wave + 1];
	const int r = blockIdx.x;
	const float* z = Z + (int64_t)r * n;
	float mx = -INFINITY;
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
	for (int i = threadIdx.x; i < n; i += 256)
		s += expf(z[i] - mx);
	s = block_sum<256>(s, s_red);
	const float lse = mx + logf(s);
	float l = 0.0f;
	for (int i = threadIdx.x; i < n; i += 256)
	{
		const float p = expf(z[i] - lse);
		P[(int64_t)r * n + i] = p;
		if (Tg)
		{
			const float t = Tg[(int64_t)r * n + i];
			if (dZ)
				dZ[(int64_t)r * n + i] = (p - t) / (float)rows;
			if (t > 0.0f)
				l -= t * (z[i] - lse);
		}
	}
	l = block_sum<256>(l, s_red);
	if (threadIdx.x == 0 && loss_rows)
		loss_rows[r] = l;
}

__global__ __launch_bounds__(256) void k_gelu_bwd(int64_t count, const float* __restrict__ Zp,
												  float* __restrict__ dA)
{
	const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if (i < count)
		dA[i] *= gelu_grad_f(Zp[i]);
}

// db[c] = sum over rows of dZ[r][c]
__global__ __launch_bounds__(256) void k_colsum(int rows, int n, const float* __restrict__ dZ, float* __restrict__ db)
{
	const int c = blockIdx.x * 256 + threadIdx.x;
	if (c >= n)
		return;
	float s = 0.0f;
	for (int r = 0; r < rows; ++r)
		s += dZ[(int64_t)r * n + c];
	db[c] = s;
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

__global__ __launch_bounds__(256) void k_sum_f(int n, const float* __restrict__ x, float* __restrict__ out)
{
	__shared__ float s_red[256 / wave + 1];
	float s = 0.0f;
	for (int i = threadIdx.x; i < n; i += 256)
		s += x[i];
	s = block_sum<256>(s, s_red);
	if (threadIdx.x == 0)
		out[0] = s;
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

void gemm(esl::ctx_t& c, int M, int N, int K, const float* A, int64_t sam, int64_t sak, const float* B, int64_t sbk,
		  int64_t sbn, const float* bias, float* C, float* Z, int epi)
{
	const dim3 grid(esl::ceil_div(N, 64), esl::ceil_div(M, 64));
	const double flops = 2.0 * M * N * K;
	esl::timed_launch(c, "k_gemm_bf16", flops, [&] {
		esl::k_gemm_bf16<<<grid, esl::gemm_block, 0, c.stream>>>(M, N, K, A, sam, sak, B, sbk, sbn, bias, C, Z, epi);
	});
}

// forward pass on rows x n_in inputs already copied to act[0]; logits in pre[last]
void forward(esl::ctx_t& c, esl_mlp* m, int rows)
{
	const int L = (int)m->dims.size() - 1;
	for (int l = 0; l < L; ++l)
	{
		const int K = m->dims[l], N = m->dims[l + 1];
		const float* X = m->act.get() + m->act_off[l];
		const float* W = m->params.get() + m->w_off[l];
		const float* b = m->params.get() + m->b_off[l];
		float* Y = m->act.get() + m->act_off[l + 1];
		float* Z = m->pre.get() + m->act_off[l + 1];
		if (l < L - 1)
			gemm(c, rows, N, K, X, K, 1, W, N, 1, b, Y, Z, esl::EPI_BIAS_GELU);
		else
			gemm(c, rows, N, K, X, K, 1, W, N, 1, b, Z, nullptr, esl::EPI_BIAS);
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
		ESL_HIP(hipMemcpyAsync(m->act.get(), d_x, (size_t)rows * m->n_in * sizeof(float), hipMemcpyDeviceToDevice,
							   c.stream));
		forward(c, m, rows);
		const int L = (int)m->dims.size() - 1;
		esl::k_softmax_xent<<<rows, 256, 0, c.stream>>>(rows, m->n_out, m->pre.get() + m->act_off[L], nullptr,
														 d_probs, nullptr, nullptr);
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
		ESL_HIP(hipMemcpyAsync(m->act.get(), d_x, (size_t)rows * m->n_in * sizeof(float), hipMemcpyDeviceToDevice,
							   c.stream));
		forward(c, m, rows);
		const int L = (int)m->dims.size() - 1;
		esl::k_softmax_xent<<<rows, 256, 0, c.stream>>>(rows, m->n_out, m->pre.get() + m->act_off[L], nullptr,
														 m->probs.get(), nullptr, nullptr);
		const int64_t cnt = (int64_t)rows * m->n_out;
		esl::k_widen<<<esl::ceil_div(cnt, 256), 256, 0, c.stream>>>(cnt, m->probs.get(), d_warm);
		ESL_HIP(hipGetLastError());
		ESL_HIP(hipStreamSynchronize(c.stream));
	});
}

// gradients of the mean cross-entropy into m->grads; returns the loss
static float mlp_backward(esl::ctx_t& c, esl_mlp* m, int rows, const float* d_x, const float* d_t)
{
	ensure_rows(m, rows);
	ESL_HIP(hipMemcpyAsync(m->act.get(), d_x, (size_t)rows * m->n_in * sizeof(float), hipMemcpyDeviceToDevice,
						   c.stream));
	forward(c, m, rows);
	const int L = (int)m->dims.size() - 1;
	esl::k_softmax_xent<<<rows, 256, 0, c.stream>>>(rows, m->n_out, m->pre.get() + m->act_off[L], d_t,
													 m->probs.get(), m->dact.get() + m->act_off[L],
													 m->loss_rows.get());
	for (int l = L - 1; l >= 0; --l)
	{
		const int K = m->dims[l], N = m->dims[l + 1];
		float* dZ = m->dact.get() + m->act_off[l + 1]; // rows x N (already multiplied by gelu')
		const float* X = m->act.get() + m->act_off[l];
		float* dW = m->grads.get() + m->w_off[l];
		float* db = m->grads.get() + m->b_off[l];
		// dW = X^T dZ  (K x N)
		gemm(c, K, N, rows, X, 1, K, dZ, N, 1, nullptr, dW, nullptr, esl::EPI_STORE);
		esl::k_colsum<<<esl::ceil_div(N, 256), 256, 0, c.stream>>>(rows, N, dZ, db);
		if (l > 0)
		{
			// dX = dZ W^T  (rows x K), then times gelu'(Z_l)
			const float* W = m->params.get() + m->w_off[l];
			float* dX = m->dact.get() + m->act_off[l];
			gemm(c, rows, K, N, dZ, N, 1, W, 1, N, nullptr, dX, nullptr, esl::EPI_STORE);
			esl::k_gelu_bwd<<<esl::ceil_div((int64_t)rows * K, 256), 256, 0, c.stream>>>(
				(int64_t)rows * K, m->pre.get() + m->act_off[l], dX);
		}
	}
	esl::k_sum_f<<<1, 256, 0, c.stream>>>(rows, m->loss_rows.get(), m->loss.get());
	ESL_HIP(hipGetLastError());
	float loss = 0.0f;
	ESL_HIP(hipMemcpyAsync(&loss, m->loss.get(), sizeof(float), hipMemcpyDeviceToHost, c.stream));
	ESL_HIP(hipStreamSynchronize(c.stream));
	return loss / rows;
}

esl_status esl_mlp_gradients(esl_ctx* ctx, esl_mlp* m, int rows, const float* d_x, const float* d_target,
							 float* host_grads, float* loss_out)
{
	return mlp_guard(ctx, [&] {
		if (!ctx || !m || !d_x || !d_target || rows < 1)
			throw esl::error(ESL_E_ARG, "null argument");
		esl::ctx_t& c = esl::ctx_impl(ctx);
		const float loss = mlp_backward(c, m, rows, d_x, d_target);
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
		const float loss = mlp_backward(c, m, rows, d_x, d_target);
		m->step++;
		const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
		const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
		esl::k_adam<<<esl::ceil_div(m->n_params, 256), 256, 0, c.stream>>>(m->n_params, m->params.get(),
																			m->grads.get(), m->m1.get(), m->m2.get(),
																			lr, b1, b2, eps, c1, c2);
		ESL_HIP(hipGetLastError());
		ESL_HIP(hipStreamSynchronize(c.stream));
		if (loss_out)
			*loss_out = loss;
	});
}

} // extern "C"
