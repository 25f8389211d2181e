// gctf_net.hip -- forward, backward and optimiser kernels of the
// defocus-proposal network, and the C-ABI of its trainer.
//
// All GEMM-shaped work (conv2-4 as im2col GEMMs, fc1, and their weight and
// input gradients) goes through one MFMA kernel, k_gemm: 64 x 64 output
// tiles, 4 waves of 32 x 32, K staged through LDS 16 at a time, operands
// read with arbitrary element strides (so transposes cost nothing),
// converted to BF16 for v_mfma_f32_32x32x16_bf16 or kept in FP32 for
// v_mfma_f32_32x32x2f32.  Long K (weight gradients over batch x length)
// is split over workgroups and reduced in a fixed order, so training is
// deterministic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gctf_hip.h"
#include "gctf_net.hpp"

namespace gctf {

typedef __bf16 nbf16x8 __attribute__((ext_vector_type(8)));
typedef float nfloatx16 __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------------------
// GEMM: C[m][n] = sum_k A(m, k) B(k, n) (+ beta C)
// ---------------------------------------------------------------------------
struct GemmArgs {
	int M, N, K;
	const float* A; long long sam, sak;
	const float* B; long long sbk, sbn;
	float* C; long long scm, scn;
	float beta;      // split == 1 only
	int ksplit;      // K range of one split (multiple of 16)
	float* part;     // [split][M][N] when gridDim.z > 1
};

template <bool BF16>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g)
{
	__shared__ float sA[64][17];
	__shared__ float sB[16][65];
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int wm = wave >> 1, wn = wave & 1;
	const int row = lane & 31, h = lane >> 5;
	const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
	const int kb = blockIdx.z * g.ksplit, ke = min(g.K, kb + g.ksplit);
	nfloatx16 acc;
#pragma unroll
	for (int r = 0; r < 16; ++r) acc[r] = 0.f;
	for (int k0 = kb; k0 < ke; k0 += 16) {
		for (int i = tid; i < 1024; i += 256) {
			const int r = i >> 4, kk = i & 15, m = m0 + r, k = k0 + kk;
			sA[r][kk] = (m < g.M && k < ke) ? g.A[(long long)m * g.sam + (long long)k * g.sak] : 0.f;
		}
		for (int i = tid; i < 1024; i += 256) {
			const int kk = i >> 6, c = i & 63, n = n0 + c, k = k0 + kk;
			sB[kk][c] = (n < g.N && k < ke) ? g.B[(long long)k * g.sbk + (long long)n * g.sbn] : 0.f;
		}
		__syncthreads();
		if constexpr (BF16) {
			nbf16x8 a, b;
#pragma unroll
			for (int j = 0; j < 8; ++j) {
				a[j] = (__bf16)sA[wm * 32 + row][8 * h + j];
				b[j] = (__bf16)sB[8 * h + j][wn * 32 + row];
			}
			acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
		} else {
#pragma unroll
			for (int kk = 0; kk < 16; kk += 2)
				acc = __builtin_amdgcn_mfma_f32_32x32x2f32(sA[wm * 32 + row][kk + h], sB[kk + h][wn * 32 + row],
					acc, 0, 0, 0);
		}
		__syncthreads();
	}
#pragma unroll
	for (int r = 0; r < 16; ++r) {
		const int m = m0 + wm * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
		const int n = n0 + wn * 32 + row;
		if (m >= g.M || n >= g.N) continue;
		if (gridDim.z > 1) {
			g.part[((long long)blockIdx.z * g.M + m) * g.N + n] = acc[r];
		} else {
			const long long ci = (long long)m * g.scm + (long long)n * g.scn;
			g.C[ci] = g.beta != 0.f ? acc[r] + g.beta * g.C[ci] : acc[r];
		}
	}
}

__global__ __launch_bounds__(256) void k_splitk_reduce(const float* __restrict__ part, int splits, int M, int N,
	float* __restrict__ C, long long scm, long long scn, float beta)
{
	const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
	if (i >= (long long)M * N) return;
	const int m = (int)(i / N), n = (int)(i % N);
	float s = 0.f;
	for (int z = 0; z < splits; ++z) s += part[(long long)z * M * N + i];
	const long long ci = (long long)m * scm + (long long)n * scn;
	C[ci] = beta != 0.f ? s + beta * C[ci] : s;
}

// ---------------------------------------------------------------------------
// layer kernels
// ---------------------------------------------------------------------------

// Input: bins [kNetBin0, 256) z-scored per spectrum, lower bins zero.
// One workgroup per spectrum, thread = bin.
__global__ __launch_bounds__(256) void k_net_input(const float* __restrict__ radial, int ld, float* __restrict__ x0)
{
	__shared__ double r1[256], r2[256];
	const int b = blockIdx.x, t = threadIdx.x;
	const float v = radial[(long long)b * ld + t];
	const bool use = t >= kNetBin0;
	r1[t] = use ? (double)v : 0.0;
	r2[t] = use ? (double)v * v : 0.0;
	__syncthreads();
	for (int s = 128; s > 0; s >>= 1) {
		if (t < s) { r1[t] += r1[t + s]; r2[t] += r2[t + s]; }
		__syncthreads();
	}
	const double cnt = (double)(kNetBins - kNetBin0);
	const double mean = r1[0] / cnt;
	const double var = fmax(r2[0] / cnt - mean * mean, 1e-24);
	x0[(long long)b * kNetBins + t] = use ? (float)(((double)v - mean) / sqrt(var)) : 0.f;
}

// conv1 (1 -> 64, 7 taps) + bias + ReLU + pool.  Thread = (channel, batch,
// output pair); y1 keeps the post-ReLU values for the backward pass.
__global__ __launch_bounds__(256) void k_conv1_fwd(const float* __restrict__ x0, const float* __restrict__ prm,
	int B, float* __restrict__ y1, float* __restrict__ a1)
{
	const int c = blockIdx.y;
	const int i = blockIdx.x * 256 + threadIdx.x;   // over B * 128 pairs
	if (i >= B * (kNetBins / 2)) return;
	const int b = i / (kNetBins / 2), p = i % (kNetBins / 2);
	const float* x = x0 + (long long)b * kNetBins;
	const float* w = prm + kNetW1 + c * kNetTaps;
	float y[2];
#pragma unroll
	for (int q = 0; q < 2; ++q) {
		const int l = 2 * p + q;
		float z = prm[kNetB1 + c];
#pragma unroll
		for (int t = 0; t < kNetTaps; ++t) {
			const int j = l + t - 3;
			if (j >= 0 && j < kNetBins) z = fmaf(w[t], x[j], z);
		}
		y[q] = fmaxf(z, 0.f);
	}
	const long long BL = (long long)B * kNetBins;
	y1[c * BL + (long long)b * kNetBins + 2 * p] = y[0];
	y1[c * BL + (long long)b * kNetBins + 2 * p + 1] = y[1];
	a1[c * (BL / 2) + (long long)b * (kNetBins / 2) + p] = 0.5f * (y[0] + y[1]);
}

// col[(ci * 7 + t)][b * L + l] = a[ci][b * L + l + t - 3] (zero outside the row)
__global__ __launch_bounds__(256) void k_im2col(const float* __restrict__ a, int B, int L, float* __restrict__ col)
{
	const int r = blockIdx.y;   // ci * 7 + t
	const int ci = r / kNetTaps, t = r % kNetTaps;
	const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
	const long long BL = (long long)B * L;
	if (q >= BL) return;
	const int l = (int)(q % L);
	const int j = l + t - 3;
	col[r * BL + q] = (j >= 0 && j < L) ? a[ci * BL + q + (t - 3)] : 0.f;
}

// z[c][:] += bias, ReLU in place, then mean-pool pairs into a.
__global__ __launch_bounds__(256) void k_bias_relu_pool(float* __restrict__ z, const float* __restrict__ bias,
	long long BL, float* __restrict__ a)
{
	const int c = blockIdx.y;
	const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
	if (p >= BL / 2) return;
	float* zr = z + c * BL;
	const float y0 = fmaxf(zr[2 * p] + bias[c], 0.f);
	const float y1 = fmaxf(zr[2 * p + 1] + bias[c], 0.f);
	zr[2 * p] = y0;
	zr[2 * p + 1] = y1;
	a[c * (BL / 2) + p] = 0.5f * (y0 + y1);
}

// f[c * 16 + l][b] = a4[c][b * 16 + l]
__global__ __launch_bounds__(256) void k_flatten(const float* __restrict__ a4, int B, float* __restrict__ f)
{
	const int i = blockIdx.x * 256 + threadIdx.x;
	if (i >= kNetFlat * B) return;
	const int fi = i / B, b = i % B;
	const int c = fi / kNetL4, l = fi % kNetL4;
	f[i] = a4[(long long)c * B * kNetL4 + b * kNetL4 + l];
}

__global__ __launch_bounds__(256) void k_unflatten(const float* __restrict__ f, int B, float* __restrict__ a4)
{
	const int i = blockIdx.x * 256 + threadIdx.x;
	if (i >= kNetFlat * B) return;
	const int fi = i / B, b = i % B;
	const int c = fi / kNetL4, l = fi % kNetL4;
	a4[(long long)c * B * kNetL4 + b * kNetL4 + l] = f[i];
}

// h[j][b] = ReLU(h[j][b] + bias[j])
__global__ __launch_bounds__(256) void k_bias_relu(float* __restrict__ h, const float* __restrict__ bias, int B)
{
	const int i = blockIdx.x * 256 + threadIdx.x;
	if (i >= kNetHidden * B) return;
	h[i] = fmaxf(h[i] + bias[i / B], 0.f);
}

// o[b][k] = b6[k] + sum_j W6[k][j] h[j][b]
__global__ __launch_bounds__(256) void k_fc2(const float* __restrict__ h, const float* __restrict__ prm, int B,
	float* __restrict__ o)
{
	const int i = blockIdx.x * 256 + threadIdx.x;
	if (i >= B * kNetOut) return;
	const int b = i / kNetOut, k = i % kNetOut;
	const float* w = prm + kNetW6 + k * kNetHidden;
	float s = prm[kNetB6 + k];
	for (int j = 0; j < kNetHidden; ++j) s = fmaf(w[j], h[(long long)j * B + b], s);
	o[i] = s;
}

// Loss per spectrum and its gradient (divided by B).  target[b] =
// (mean defocus A, astigmatism A, azimuth deg, confidence label).
//   100 (o0 - ln(df / 1e4))^2 + 0.25 (o1 - ast / 1e3)^2
//   + 0.05 (1 - cos 2 (o2 - az)) + 0.5 w BCE(sigmoid(o3), c)
// The label is measured on the micrograph (whether the learned path's
// narrowed search ended within 20 A of the full search): > 0 gives c = 1,
// < 0 gives c = 0, and 0 (unlabelled) drops the confidence term (w = 0).
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ o, const float* __restrict__ tg, int B,
	float* __restrict__ dO, double* __restrict__ lpart)
{
	const int b = blockIdx.x * 256 + threadIdx.x;
	if (b >= B) return;
	const float* ob = o + b * kNetOut;
	const float* tb = tg + b * 4;
	const float y0 = logf(tb[0] * 1.0e-4f), y1 = tb[1] * 1.0e-3f;
	const float az = tb[2] * 0.017453292519943295f;
	const float e0 = ob[0] - y0, e1 = ob[1] - y1;
	const float d2 = 2.0f * (ob[2] - az);
	const float c = tb[3] > 0.0f ? 1.0f : 0.0f;
	const float w = tb[3] != 0.0f ? 1.0f : 0.0f;
	const float x = ob[3];
	// softplus forms of the log-sigmoid terms
	const float sp_pos = fmaxf(x, 0.f) + log1pf(expf(-fabsf(x)));   // -log(1 - p)
	const float sp_neg = sp_pos - x;                                 // -log p
	const float bce = c * sp_neg + (1.0f - c) * sp_pos;
	const float p = 1.0f / (1.0f + expf(-x));
	const float L = 100.0f * e0 * e0 + 0.25f * e1 * e1 + 0.05f * (1.0f - cosf(d2)) + 0.5f * w * bce;
	const float inv = 1.0f / (float)B;
	dO[b * kNetOut + 0] = 200.0f * e0 * inv;
	dO[b * kNetOut + 1] = 0.5f * e1 * inv;
	dO[b * kNetOut + 2] = 0.1f * sinf(d2) * inv;
	dO[b * kNetOut + 3] = 0.5f * w * (p - c) * inv;
	lpart[b] = (double)L;
}

__global__ __launch_bounds__(256) void k_loss_sum(const double* __restrict__ lpart, int B, double* __restrict__ out)
{
	__shared__ double r[256];
	double s = 0.0;
	for (int i = threadIdx.x; i < B; i += 256) s += lpart[i];
	r[threadIdx.x] = s;
	__syncthreads();
	for (int o = 128; o > 0; o >>= 1) {
		if (threadIdx.x < o) r[threadIdx.x] += r[threadIdx.x + o];
		__syncthreads();
	}
	if (threadIdx.x == 0) out[0] = r[0] / (double)B;
}

// dh[j][b] = (h > 0) sum_k W6[k][j] dO[b][k]
__global__ __launch_bounds__(256) void k_fc2_dh(const float* __restrict__ dO, const float* __restrict__ h,
	const float* __restrict__ prm, int B, float* __restrict__ dh)
{
	const int i = blockIdx.x * 256 + threadIdx.x;
	if (i >= kNetHidden * B) return;
	const int j = i / B, b = i % B;
	float s = 0.f;
#pragma unroll
	for (int k = 0; k < kNetOut; ++k) s = fmaf(prm[kNetW6 + k * kNetHidden + j], dO[b * kNetOut + k], s);
	dh[i] = h[i] > 0.f ? s : 0.f;
}

// dW6[k][j] = sum_b dO[b][k] h[j][b]; db6[k] = sum_b dO[b][k]
__global__ __launch_bounds__(256) void k_fc2_dw(const float* __restrict__ dO, const float* __restrict__ h, int B,
	float* __restrict__ grad)
{
	const int i = blockIdx.x * 256 + threadIdx.x;
	if (i >= kNetOut * (kNetHidden + 1)) return;
	const int k = i / (kNetHidden + 1), j = i % (kNetHidden + 1);
	float s = 0.f;
	if (j < kNetHidden) {
		for (int b = 0; b < B; ++b) s = fmaf(dO[b * kNetOut + k], h[(long long)j * B + b], s);
		grad[kNetW6 + k * kNetHidden + j] = s;
	} else {
		for (int b = 0; b < B; ++b) s += dO[b * kNetOut + k];
		grad[kNetB6 + k] = s;
	}
}

// out[r] = sum over the row X[r][0..n)
__global__ __launch_bounds__(256) void k_rowsum(const float* __restrict__ X, long long n, float* __restrict__ out)
{
	__shared__ float red[256];
	const float* x = X + (long long)blockIdx.x * n;
	float s = 0.f;
	for (long long i = threadIdx.x; i < n; i += 256) s += x[i];
	red[threadIdx.x] = s;
	__syncthreads();
	for (int o = 128; o > 0; o >>= 1) {
		if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
		__syncthreads();
	}
	if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// dZ[c][i] = (y[c][i] > 0) 0.5 dA[c][i / 2]  (mean-pool then ReLU backward)
__global__ __launch_bounds__(256) void k_pool_relu_back(const float* __restrict__ y, const float* __restrict__ dA,
	long long BL, float* __restrict__ dZ)
{
	const int c = blockIdx.y;
	const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
	if (i >= BL) return;
	dZ[c * BL + i] = y[c * BL + i] > 0.f ? 0.5f * dA[c * (BL / 2) + i / 2] : 0.f;
}

// dA[ci][b * L + l] = sum_t dcol[ci * 7 + t][b * L + l - t + 3] (valid taps)
__global__ __launch_bounds__(256) void k_col2im(const float* __restrict__ dcol, int B, int L, float* __restrict__ dA)
{
	const int ci = blockIdx.y;
	const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
	const long long BL = (long long)B * L;
	if (q >= BL) return;
	const int l = (int)(q % L);
	float s = 0.f;
#pragma unroll
	for (int t = 0; t < kNetTaps; ++t) {
		const int j = l - t + 3;
		if (j >= 0 && j < L) s += dcol[(long long)(ci * kNetTaps + t) * BL + q - t + 3];
	}
	dA[ci * BL + q] = s;
}

// conv1 weight and bias gradients: one workgroup per output channel.
__global__ __launch_bounds__(256) void k_conv1_dw(const float* __restrict__ dZ, const float* __restrict__ x0, int B,
	float* __restrict__ grad)
{
	__shared__ float red[kNetTaps + 1][256];
	const int c = blockIdx.x;
	const long long BL = (long long)B * kNetBins;
	float s[kNetTaps + 1];
#pragma unroll
	for (int t = 0; t <= kNetTaps; ++t) s[t] = 0.f;
	for (long long q = threadIdx.x; q < BL; q += 256) {
		const float d = dZ[c * BL + q];
		const int l = (int)(q % kNetBins);
		const float* x = x0 + (q - l);
#pragma unroll
		for (int t = 0; t < kNetTaps; ++t) {
			const int j = l + t - 3;
			if (j >= 0 && j < kNetBins) s[t] = fmaf(d, x[j], s[t]);
		}
		s[kNetTaps] += d;
	}
#pragma unroll
	for (int t = 0; t <= kNetTaps; ++t) red[t][threadIdx.x] = s[t];
	__syncthreads();
	for (int o = 128; o > 0; o >>= 1) {
		if (threadIdx.x < o)
#pragma unroll
			for (int t = 0; t <= kNetTaps; ++t) red[t][threadIdx.x] += red[t][threadIdx.x + o];
		__syncthreads();
	}
	if (threadIdx.x < kNetTaps) grad[kNetW1 + c * kNetTaps + threadIdx.x] = red[threadIdx.x][0];
	if (threadIdx.x == kNetTaps) grad[kNetB1 + c] = red[kNetTaps][0];
}

// AdamW: decoupled weight decay on the weights, bias-corrected moments.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, const float* __restrict__ g,
	float* __restrict__ m, float* __restrict__ v, long long n, float lr, float b1, float b2, float eps,
	float wd, float bc1, float bc2)
{
	const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
	if (i >= n) return;
	const float gi = g[i];
	const float mi = b1 * m[i] + (1.0f - b1) * gi;
	const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
	m[i] = mi;
	v[i] = vi;
	const float upd = (mi / bc1) / (sqrtf(vi / bc2) + eps);
	p[i] = p[i] - lr * (upd + wd * p[i]);
}

// ---------------------------------------------------------------------------
// buffers and launch sequences
// ---------------------------------------------------------------------------
struct NetBuffers {
	int cap = 0;
	bool train = false;
	std::vector<void*> owned;
	float *x0 = nullptr, *y1 = nullptr, *a1 = nullptr, *col = nullptr, *y2 = nullptr, *a2 = nullptr;
	float *y3 = nullptr, *a3 = nullptr, *y4 = nullptr, *a4 = nullptr, *f = nullptr, *h = nullptr, *o = nullptr;
	// training
	float *dO = nullptr, *dh = nullptr, *dF = nullptr, *dA = nullptr, *dA2 = nullptr, *dZ = nullptr;
	float *dcol = nullptr, *grad = nullptr, *part = nullptr;
	double *lpart = nullptr, *loss = nullptr;
	long long part_cap = 0;
};

namespace {

hipError_t alloc_f(NetBuffers* b, float** p, long long n)
{
	void* q = nullptr;
	hipError_t e = hipMalloc(&q, sizeof(float) * (size_t)std::max<long long>(n, 1));
	if (e != hipSuccess) return e;
	b->owned.push_back(q);
	*p = (float*)q;
	return hipSuccess;
}

#define NETCHK(x)                                  \
	do {                                           \
		hipError_t e_ = (x);                       \
		if (e_ != hipSuccess) return e_;           \
	} while (0)

hipError_t gemm(NetBuffers* nb, bool bf16, GemmArgs g, hipStream_t s)
{
	// split long K so that the grid has enough workgroups
	const int tiles = ((g.N + 63) / 64) * ((g.M + 63) / 64);
	int splits = 1;
	if (tiles < 256 && g.K >= 128 && nb->part) {
		// at least 64 K per split, enough workgroups to cover the chip
		splits = std::min(g.K / 64, std::max(1, 512 / tiles));
		while (splits > 1 && (long long)splits * g.M * g.N > nb->part_cap) --splits;
	}
	g.ksplit = ((g.K + splits - 1) / splits + 15) / 16 * 16;
	splits = (g.K + g.ksplit - 1) / g.ksplit;
	g.part = nb->part;
	dim3 grid((g.N + 63) / 64, (g.M + 63) / 64, splits);
	if (bf16) k_gemm<true><<<grid, 256, 0, s>>>(g);
	else k_gemm<false><<<grid, 256, 0, s>>>(g);
	if (splits > 1) {
		const long long mn = (long long)g.M * g.N;
		k_splitk_reduce<<<(unsigned)((mn + 255) / 256), 256, 0, s>>>(g.part, splits, g.M, g.N, g.C, g.scm, g.scn,
			g.beta);
	}
	return hipGetLastError();
}

inline unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }

// conv layer l (2..4) forward: input a_in [64][B L], im2col, GEMM, bias,
// ReLU, pool into a_out [64][B L / 2]; y keeps the post-ReLU values.
hipError_t conv_fwd(NetBuffers* nb, bool bf16, const float* prm, long long woff, long long boff,
	const float* a_in, int B, int L, float* y, float* a_out, hipStream_t s)
{
	const long long BL = (long long)B * L;
	k_im2col<<<dim3(nblk(BL), kNetCol), 256, 0, s>>>(a_in, B, L, nb->col);
	GemmArgs g{};
	g.M = kNetCh; g.N = (int)BL; g.K = kNetCol;
	g.A = prm + woff; g.sam = kNetCol; g.sak = 1;
	g.B = nb->col; g.sbk = BL; g.sbn = 1;
	g.C = y; g.scm = BL; g.scn = 1;
	NETCHK(gemm(nb, bf16, g, s));
	k_bias_relu_pool<<<dim3(nblk(BL / 2), kNetCh), 256, 0, s>>>(y, prm + boff, BL, a_out);
	return hipGetLastError();
}

hipError_t forward_all(NetBuffers* nb, const float* prm, const float* radial, int ld, int B, bool bf16,
	hipStream_t s)
{
	k_net_input<<<B, 256, 0, s>>>(radial, ld, nb->x0);
	k_conv1_fwd<<<dim3(nblk((long long)B * kNetBins / 2), kNetCh), 256, 0, s>>>(nb->x0, prm, B, nb->y1, nb->a1);
	NETCHK(conv_fwd(nb, bf16, prm, kNetW2, kNetB2, nb->a1, B, kNetBins / 2, nb->y2, nb->a2, s));
	NETCHK(conv_fwd(nb, bf16, prm, kNetW3, kNetB3, nb->a2, B, kNetBins / 4, nb->y3, nb->a3, s));
	NETCHK(conv_fwd(nb, bf16, prm, kNetW4, kNetB4, nb->a3, B, kNetBins / 8, nb->y4, nb->a4, s));
	k_flatten<<<nblk((long long)kNetFlat * B), 256, 0, s>>>(nb->a4, B, nb->f);
	GemmArgs g{};
	g.M = kNetHidden; g.N = B; g.K = kNetFlat;
	g.A = prm + kNetW5; g.sam = kNetFlat; g.sak = 1;
	g.B = nb->f; g.sbk = B; g.sbn = 1;
	g.C = nb->h; g.scm = B; g.scn = 1;
	NETCHK(gemm(nb, bf16, g, s));
	k_bias_relu<<<nblk((long long)kNetHidden * B), 256, 0, s>>>(nb->h, prm + kNetB5, B);
	k_fc2<<<nblk((long long)B * kNetOut), 256, 0, s>>>(nb->h, prm, B, nb->o);
	return hipGetLastError();
}

// conv layer backward: dA_out [64][B L / 2] -> grads of W, b and (when
// dA_in) the layer input.
hipError_t conv_back(NetBuffers* nb, bool bf16, const float* prm, long long woff, long long boff,
	const float* a_in, const float* y, const float* dA_out, int B, int L, float* dA_in, hipStream_t s)
{
	const long long BL = (long long)B * L;
	k_pool_relu_back<<<dim3(nblk(BL), kNetCh), 256, 0, s>>>(y, dA_out, BL, nb->dZ);
	// the forward im2col of this layer's input (the column buffer is shared)
	k_im2col<<<dim3(nblk(BL), kNetCol), 256, 0, s>>>(a_in, B, L, nb->col);
	GemmArgs g{};
	g.M = kNetCh; g.N = kNetCol; g.K = (int)BL;
	g.A = nb->dZ; g.sam = BL; g.sak = 1;
	g.B = nb->col; g.sbk = 1; g.sbn = BL;
	g.C = nb->grad + woff; g.scm = kNetCol; g.scn = 1;
	NETCHK(gemm(nb, bf16, g, s));
	k_rowsum<<<kNetCh, 256, 0, s>>>(nb->dZ, BL, nb->grad + boff);
	if (dA_in) {
		GemmArgs d{};
		d.M = kNetCol; d.N = (int)BL; d.K = kNetCh;
		d.A = prm + woff; d.sam = 1; d.sak = kNetCol;
		d.B = nb->dZ; d.sbk = BL; d.sbn = 1;
		d.C = nb->dcol; d.scm = BL; d.scn = 1;
		NETCHK(gemm(nb, bf16, d, s));
		k_col2im<<<dim3(nblk(BL), kNetCh), 256, 0, s>>>(nb->dcol, B, L, dA_in);
	}
	return hipGetLastError();
}

hipError_t backward_all(NetBuffers* nb, const float* prm, const float* target, int B, bool bf16, hipStream_t s)
{
	k_loss<<<nblk(B), 256, 0, s>>>(nb->o, target, B, nb->dO, nb->lpart);
	k_loss_sum<<<1, 256, 0, s>>>(nb->lpart, B, nb->loss);
	k_fc2_dw<<<nblk(kNetOut * (kNetHidden + 1)), 256, 0, s>>>(nb->dO, nb->h, B, nb->grad);
	k_fc2_dh<<<nblk((long long)kNetHidden * B), 256, 0, s>>>(nb->dO, nb->h, prm, B, nb->dh);
	GemmArgs g{};
	g.M = kNetHidden; g.N = kNetFlat; g.K = B;
	g.A = nb->dh; g.sam = B; g.sak = 1;
	g.B = nb->f; g.sbk = 1; g.sbn = B;
	g.C = nb->grad + kNetW5; g.scm = kNetFlat; g.scn = 1;
	NETCHK(gemm(nb, bf16, g, s));
	k_rowsum<<<kNetHidden, 256, 0, s>>>(nb->dh, B, nb->grad + kNetB5);
	GemmArgs d{};
	d.M = kNetFlat; d.N = B; d.K = kNetHidden;
	d.A = prm + kNetW5; d.sam = 1; d.sak = kNetFlat;
	d.B = nb->dh; d.sbk = B; d.sbn = 1;
	d.C = nb->dF; d.scm = B; d.scn = 1;
	NETCHK(gemm(nb, bf16, d, s));
	k_unflatten<<<nblk((long long)kNetFlat * B), 256, 0, s>>>(nb->dF, B, nb->dA);
	// conv4: dA (pooled, [64][B 16]) -> dA2 ([64][B 32])
	NETCHK(conv_back(nb, bf16, prm, kNetW4, kNetB4, nb->a3, nb->y4, nb->dA, B, kNetBins / 8, nb->dA2, s));
	NETCHK(conv_back(nb, bf16, prm, kNetW3, kNetB3, nb->a2, nb->y3, nb->dA2, B, kNetBins / 4, nb->dA, s));
	NETCHK(conv_back(nb, bf16, prm, kNetW2, kNetB2, nb->a1, nb->y2, nb->dA, B, kNetBins / 2, nb->dA2, s));
	// conv1
	const long long BL = (long long)B * kNetBins;
	k_pool_relu_back<<<dim3(nblk(BL), kNetCh), 256, 0, s>>>(nb->y1, nb->dA2, BL, nb->dZ);
	k_conv1_dw<<<kNetCh, 256, 0, s>>>(nb->dZ, nb->x0, B, nb->grad);
	return hipGetLastError();
}

}  // namespace

hipError_t net_buffers_create(int max_batch, bool train, NetBuffers** out)
{
	*out = nullptr;
	if (max_batch < 1) return hipErrorInvalidValue;
	NetBuffers* nb = new NetBuffers;
	nb->cap = max_batch;
	nb->train = train;
	const long long B = max_batch, L = kNetBins;
	hipError_t e = hipSuccess;
	auto A = [&](float** p, long long n) { if (e == hipSuccess) e = alloc_f(nb, p, n); };
	A(&nb->x0, B * L);
	A(&nb->y1, kNetCh * B * L);
	A(&nb->a1, kNetCh * B * L / 2);
	A(&nb->col, (long long)kNetCol * B * L / 2);
	A(&nb->y2, kNetCh * B * L / 2);
	A(&nb->a2, kNetCh * B * L / 4);
	A(&nb->y3, kNetCh * B * L / 4);
	A(&nb->a3, kNetCh * B * L / 8);
	A(&nb->y4, kNetCh * B * L / 8);
	A(&nb->a4, kNetCh * B * L / 16);
	A(&nb->f, kNetFlat * B);
	A(&nb->h, kNetHidden * B);
	A(&nb->o, kNetOut * B);
	// split-K partials of the GEMMs (forward and backward)
	nb->part_cap = std::max<long long>(1LL << 20, (long long)kNetCh * B * L * 2);
	A(&nb->part, nb->part_cap);
	if (train) {
		A(&nb->dO, kNetOut * B);
		A(&nb->dh, kNetHidden * B);
		A(&nb->dF, kNetFlat * B);
		A(&nb->dA, kNetCh * B * L / 2);
		A(&nb->dA2, kNetCh * B * L / 2);
		A(&nb->dZ, kNetCh * B * L);
		A(&nb->dcol, (long long)kNetCol * B * L / 2);
		A(&nb->grad, kNetParams);
		float* lp = nullptr;
		A(&lp, 2 * B + 2);
		if (e == hipSuccess) {
			nb->lpart = (double*)lp;
			nb->loss = nb->lpart + B;
		}
	}
	if (e != hipSuccess) {
		net_buffers_destroy(nb);
		return e;
	}
	*out = nb;
	return hipSuccess;
}

void net_buffers_destroy(NetBuffers* b)
{
	if (!b) return;
	for (void* p : b->owned) (void)hipFree(p);
	delete b;
}

int net_buffers_capacity(const NetBuffers* b) { return b ? b->cap : 0; }

hipError_t net_forward(NetBuffers* b, const float* params, const float* radial, int ld, int n, bool bf16,
	float* out, hipStream_t s)
{
	if (!b || n < 1 || n > b->cap || ld < kNetBins) return hipErrorInvalidValue;
	NETCHK(forward_all(b, params, radial, ld, n, bf16, s));
	return hipMemcpyAsync(out, b->o, sizeof(float) * kNetOut * n, hipMemcpyDeviceToDevice, s);
}

}  // namespace gctf

// ---------------------------------------------------------------------------
// trainer C-ABI
// ---------------------------------------------------------------------------
using namespace gctf;

struct gctf_net {
	int device = 0;
	hipStream_t stream = nullptr;
	NetBuffers* nb = nullptr;
	float* params = nullptr;
	float* m = nullptr;
	float* v = nullptr;
	long long step = 0;
	std::string err;
};

namespace {
int net_fail(gctf_net* n, int code, const char* what)
{
	if (n) n->err = what;
	return code;
}
#define NETHIP(n, call)                                                           \
	do {                                                                          \
		hipError_t e_ = (call);                                                   \
		if (e_ != hipSuccess) return net_fail(n, GCTF_ERR_HIP, hipGetErrorString(e_)); \
	} while (0)
}  // namespace

long long gctf_net_param_count(void) { return kNetParams; }

int gctf_net_create(int device, int max_batch, gctf_net** out)
{
	if (!out || max_batch < 1) return GCTF_ERR_ARG;
	*out = nullptr;
	if (hipSetDevice(device) != hipSuccess) return GCTF_ERR_HIP;
	gctf_net* n = new gctf_net;
	n->device = device;
	if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) { delete n; return GCTF_ERR_HIP; }
	if (net_buffers_create(max_batch, true, &n->nb) != hipSuccess
		|| hipMalloc(&n->params, sizeof(float) * kNetParams) != hipSuccess
		|| hipMalloc(&n->m, sizeof(float) * kNetParams) != hipSuccess
		|| hipMalloc(&n->v, sizeof(float) * kNetParams) != hipSuccess) {
		gctf_net_destroy(n);
		return GCTF_ERR_NOMEM;
	}
	(void)hipMemset(n->params, 0, sizeof(float) * kNetParams);
	(void)hipMemset(n->m, 0, sizeof(float) * kNetParams);
	(void)hipMemset(n->v, 0, sizeof(float) * kNetParams);
	if (hipDeviceSynchronize() != hipSuccess) { gctf_net_destroy(n); return GCTF_ERR_HIP; }
	*out = n;
	return GCTF_OK;
}

void gctf_net_destroy(gctf_net* n)
{
	if (!n) return;
	(void)hipSetDevice(n->device);
	if (n->stream) (void)hipStreamSynchronize(n->stream);
	net_buffers_destroy(n->nb);
	if (n->params) (void)hipFree(n->params);
	if (n->m) (void)hipFree(n->m);
	if (n->v) (void)hipFree(n->v);
	if (n->stream) (void)hipStreamDestroy(n->stream);
	delete n;
}

const char* gctf_net_error_string(const gctf_net* n) { return n ? n->err.c_str() : "null network"; }

int gctf_net_set_params(gctf_net* n, const float* host, long long count)
{
	if (!n || !host || count != kNetParams) return net_fail(n, GCTF_ERR_ARG, "parameter count");
	NETHIP(n, hipSetDevice(n->device));
	NETHIP(n, hipMemcpy(n->params, host, sizeof(float) * kNetParams, hipMemcpyHostToDevice));
	return GCTF_OK;
}

int gctf_net_get_params(gctf_net* n, float* host, long long count)
{
	if (!n || !host || count != kNetParams) return net_fail(n, GCTF_ERR_ARG, "parameter count");
	NETHIP(n, hipSetDevice(n->device));
	NETHIP(n, hipStreamSynchronize(n->stream));
	NETHIP(n, hipMemcpy(host, n->params, sizeof(float) * kNetParams, hipMemcpyDeviceToHost));
	return GCTF_OK;
}

int gctf_net_reset_optimizer(gctf_net* n)
{
	if (!n) return GCTF_ERR_ARG;
	NETHIP(n, hipSetDevice(n->device));
	NETHIP(n, hipMemsetAsync(n->m, 0, sizeof(float) * kNetParams, n->stream));
	NETHIP(n, hipMemsetAsync(n->v, 0, sizeof(float) * kNetParams, n->stream));
	n->step = 0;
	return GCTF_OK;
}

int gctf_net_forward(gctf_net* n, const float* d_radial, int ld, int count, int bf16, float* host_out)
{
	if (!n || !d_radial || !host_out) return net_fail(n, GCTF_ERR_ARG, "null argument");
	if (count < 1 || count > n->nb->cap) return net_fail(n, GCTF_ERR_ARG, "batch exceeds the network capacity");
	if (ld < kNetBins) return net_fail(n, GCTF_ERR_ARG, "radial row shorter than 256 bins");
	NETHIP(n, hipSetDevice(n->device));
	NETHIP(n, forward_all(n->nb, n->params, d_radial, ld, count, bf16 != 0, n->stream));
	NETHIP(n, hipMemcpyAsync(host_out, n->nb->o, sizeof(float) * kNetOut * count, hipMemcpyDeviceToHost,
		n->stream));
	NETHIP(n, hipStreamSynchronize(n->stream));
	return GCTF_OK;
}

int gctf_net_loss_grad(gctf_net* n, const float* d_radial, int ld, const float* d_target, int count, int bf16,
	double* loss, float* host_grad)
{
	if (!n || !d_radial || !d_target) return net_fail(n, GCTF_ERR_ARG, "null argument");
	if (count < 1 || count > n->nb->cap) return net_fail(n, GCTF_ERR_ARG, "batch exceeds the network capacity");
	if (ld < kNetBins) return net_fail(n, GCTF_ERR_ARG, "radial row shorter than 256 bins");
	NETHIP(n, hipSetDevice(n->device));
	NETHIP(n, forward_all(n->nb, n->params, d_radial, ld, count, bf16 != 0, n->stream));
	NETHIP(n, backward_all(n->nb, n->params, d_target, count, bf16 != 0, n->stream));
	if (loss) NETHIP(n, hipMemcpyAsync(loss, n->nb->loss, sizeof(double), hipMemcpyDeviceToHost, n->stream));
	if (host_grad)
		NETHIP(n, hipMemcpyAsync(host_grad, n->nb->grad, sizeof(float) * kNetParams, hipMemcpyDeviceToHost,
			n->stream));
	if (loss || host_grad) NETHIP(n, hipStreamSynchronize(n->stream));
	return GCTF_OK;
}

int gctf_net_adam_step(gctf_net* n, float lr, float beta1, float beta2, float eps, float weight_decay)
{
	if (!n) return GCTF_ERR_ARG;
	NETHIP(n, hipSetDevice(n->device));
	n->step += 1;
	const float bc1 = (float)(1.0 - std::pow((double)beta1, (double)n->step));
	const float bc2 = (float)(1.0 - std::pow((double)beta2, (double)n->step));
	k_adam<<<(unsigned)((kNetParams + 255) / 256), 256, 0, n->stream>>>(n->params, n->nb->grad, n->m, n->v,
		kNetParams, lr, beta1, beta2, eps, weight_decay, bc1, bc2);
	NETHIP(n, hipGetLastError());
	return GCTF_OK;
}

int gctf_net_synchronize(gctf_net* n)
{
	if (!n) return GCTF_ERR_ARG;
	NETHIP(n, hipStreamSynchronize(n->stream));
	return GCTF_OK;
}
