// k_grid.hpp -- exhaustive defocus x astigmatism x azimuth grid search as a
// BF16 GEMM on MFMA (v_mfma_f32_32x32x16_bf16), FP32 accumulation.
//
//   numerator[m, n] = sum_k c_m[k] * S_n[k]
//   c_m[k] = (|CTF_m(k)| - 0.5) * exp(-16 s_k^2)   (the GCC2D weighting)
//   CC[m, n] = numerator / sqrt(E_m F_n),  E_m = sum c_m^2, F_n = sum S_n^2
//
// m runs over the candidate grid (shared by the whole batch), k over the
// ring pixels, n over the micrographs.  The A operand (candidates x pixels)
// would be tens of GB, so each wave generates its 32 x 16 fragment of CTF
// values in registers right before the MFMA that consumes it.  The best 16
// candidates per micrograph are then re-scored in FP32/FP64 with the exact
// evaluator of the 2D search.
//
// MFMA operand layout (v_mfma_f32_32x32x16_bf16, wave64): lane l holds
// A[l % 32][8 (l / 32) + j] and B[8 (l / 32) + j][l % 32], j = 0..7; the
// result register r of lane l is D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
#pragma once
#include "gctf_internal.hpp"
#include "k_search.hpp"

namespace gctf {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kGridTopK = 16;

// Candidate grid: m = (i_df * n_ast + i_ast) * n_az + i_az.
struct GridSpec {
	float df_lo, df_step;   // A
	float ast_lo, ast_step; // ratio
	float az_lo, az_step;   // degree
	int n_df, n_ast, n_az;
	int m;                  // n_df * n_ast * n_az
};

__device__ __forceinline__ void grid_candidate(const GridSpec& g, int m, float& dfm, float& ast, float& az)
{
	const int iaz = m % g.n_az;
	const int rest = m / g.n_az;
	const int iast = rest % g.n_ast, idf = rest / g.n_ast;
	dfm = g.df_lo + (float)idf * g.df_step;
	ast = g.ast_lo + (float)iast * g.ast_step;
	az = g.az_lo + (float)iaz * g.az_step;
}

// Ring spectra of the batch in the B-operand layout:
// bt[(k / 16) * np + n][k % 16] as bf16, zero for k >= K or n >= nimg;
// fn[n] = sum over k of the bf16 values squared (FP64).
__global__ __launch_bounds__(256) void k_grid_pack(const float* __restrict__ ring_spec, int K, int k16,
	int nimg, int np, __bf16* __restrict__ bt, double* __restrict__ fn)
{
	__shared__ double red[256];
	const int n = blockIdx.x;
	double acc = 0.0;
	for (int k = threadIdx.x; k < k16; k += 256) {
		float v = (n < nimg && k < K) ? ring_spec[(long long)n * K + k] : 0.f;
		__bf16 h = (__bf16)v;
		float r = (float)h;
		acc += (double)r * r;
		bt[((long long)(k >> 4) * np + n) * 16 + (k & 15)] = h;
	}
	red[threadIdx.x] = acc;
	__syncthreads();
	for (int s = 128; s > 0; s >>= 1) {
		if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
		__syncthreads();
	}
	if (threadIdx.x == 0) fn[n] = red[0];
}

// The GEMM.  One workgroup = 4 waves = 128 candidates x (32 * NT)
// micrographs; grid (ceil(M / 128)).  cc: [M][np] floats (m-major).
template <int NT>
__global__ __launch_bounds__(256) void k_grid_gemm(GridSpec g, CtfConsts cc, RingTables ring, int k16,
	const __bf16* __restrict__ bt, const double* __restrict__ fn, int np, float ext_phase_deg,
	float* __restrict__ out)
{
	__shared__ float s_e[4][32];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int half = lane >> 5, row = lane & 31;
	const int m = blockIdx.x * 128 + wave * 32 + row;
	const bool live = m < g.m;
	// this lane's candidate (GCalcCTF2D::DoIt form, as in the 2D search)
	float mean = 0.f, sig = 0.f, ca = 1.f, sa = 0.f, ph0 = 0.f;
	if (live) {
		float dfm, ast, az;
		grid_candidate(g, m, dfm, ast, az);
		CandCtf c = cand_ctf(cc, dfm, ast, az, ext_phase_deg);
		mean = c.mean; sig = c.sig; ca = c.ca; sa = c.sa; ph0 = c.ph0;
	}
	floatx16 acc[NT];
#pragma unroll
	for (int t = 0; t < NT; ++t)
#pragma unroll
		for (int r = 0; r < 16; ++r) acc[t][r] = 0.f;
	float e = 0.f;
	const int steps = k16 >> 4;
	for (int st = 0; st < steps; ++st) {
		const int k0 = st * 16 + half * 8;   // ring tables are zero-padded to k16
		float ra[8], rb[8], rc[8], rs[8], rw[8];
		*reinterpret_cast<float4*>(ra) = *reinterpret_cast<const float4*>(ring.a + k0);
		*reinterpret_cast<float4*>(ra + 4) = *reinterpret_cast<const float4*>(ring.a + k0 + 4);
		*reinterpret_cast<float4*>(rb) = *reinterpret_cast<const float4*>(ring.b + k0);
		*reinterpret_cast<float4*>(rb + 4) = *reinterpret_cast<const float4*>(ring.b + k0 + 4);
		*reinterpret_cast<float4*>(rc) = *reinterpret_cast<const float4*>(ring.c2 + k0);
		*reinterpret_cast<float4*>(rc + 4) = *reinterpret_cast<const float4*>(ring.c2 + k0 + 4);
		*reinterpret_cast<float4*>(rs) = *reinterpret_cast<const float4*>(ring.s2 + k0);
		*reinterpret_cast<float4*>(rs + 4) = *reinterpret_cast<const float4*>(ring.s2 + k0 + 4);
		*reinterpret_cast<float4*>(rw) = *reinterpret_cast<const float4*>(ring.w + k0);
		*reinterpret_cast<float4*>(rw + 4) = *reinterpret_cast<const float4*>(ring.w + k0 + 4);
		bf16x8 a;
#pragma unroll
		for (int j = 0; j < 8; ++j) {
			const float gg = fmaf(rc[j], ca, rs[j] * sa);
			const float df = fmaf(sig, gg, mean);
			const float u = fmaf(ra[j], df, ph0 - rb[j]);
			// |sin(pi u)| = |sin(2 pi (u / 2))|; the hardware sine takes revolutions
			const float sv = fabsf(__builtin_amdgcn_sinf(0.5f * u));
			const float cv = live ? (sv - 0.5f) * rw[j] : 0.f;
			const __bf16 h = (__bf16)cv;
			const float hr = (float)h;
			e = fmaf(hr, hr, e);
			a[j] = h;
		}
		const __bf16* bp = bt + ((long long)st * np + row) * 16 + half * 8;
#pragma unroll
		for (int t = 0; t < NT; ++t) {
			const bf16x8 b = *reinterpret_cast<const bf16x8*>(bp + (long long)t * 32 * 16);
			acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[t], 0, 0, 0);
		}
	}
	// E_m: the two k-halves of a row live in lanes l and l ^ 32
	e += __shfl_xor(e, 32);
	if (half == 0) s_e[wave][row] = e;
	__syncthreads();
	const int mrow0 = blockIdx.x * 128 + wave * 32;
#pragma unroll
	for (int t = 0; t < NT; ++t) {
		const int col = t * 32 + row;
		const double f = fn[col];
#pragma unroll
		for (int r = 0; r < 16; ++r) {
			const int drow = 8 * (r >> 2) + 4 * half + (r & 3);
			const int mm = mrow0 + drow;
			if (mm >= g.m) continue;
			const double d = sqrt((double)s_e[wave][drow] * f);
			out[(long long)mm * np + col] = d > 0.0 ? (float)((double)acc[t][r] / d) : 0.f;
		}
	}
}

// Per-chunk top-k: thread = micrograph column, rows [r0, r1) of cc.
// part[chunk][col][k] = (cc, m), sorted descending; ties keep the smaller m.
__global__ __launch_bounds__(256) void k_grid_topk(const float* __restrict__ cc, int m_total, int np,
	int nimg, int rows, float2* __restrict__ part)
{
	const int col = threadIdx.x;
	if (col >= nimg) return;
	const int r0 = blockIdx.x * rows, r1 = min(m_total, r0 + rows);
	float v[kGridTopK];
	int ix[kGridTopK];
#pragma unroll
	for (int i = 0; i < kGridTopK; ++i) { v[i] = -3.0e38f; ix[i] = 0x7fffffff; }
	for (int mm = r0; mm < r1; ++mm) {
		const float x = cc[(long long)mm * np + col];
		if (!(x > v[kGridTopK - 1])) continue;
		// insert (strictly greater values move up; equal values keep order)
		float cv = x;
		int ci = mm;
#pragma unroll
		for (int i = 0; i < kGridTopK; ++i) {
			if (cv > v[i]) {
				float tv = v[i]; int ti = ix[i];
				v[i] = cv; ix[i] = ci;
				cv = tv; ci = ti;
			}
		}
	}
	float2* out = part + ((long long)blockIdx.x * np + col) * kGridTopK;
#pragma unroll
	for (int i = 0; i < kGridTopK; ++i) out[i] = make_float2(v[i], __int_as_float(ix[i]));
}

// (cc, m) ordering of the top-k lists: larger cc first, then smaller m.
__device__ __forceinline__ void topk_insert(float* v, int* ix, float cv, int ci)
{
#pragma unroll
	for (int i = 0; i < kGridTopK; ++i) {
		if (cv > v[i] || (cv == v[i] && ci < ix[i])) {
			float tv = v[i]; int ti = ix[i];
			v[i] = cv; ix[i] = ci;
			cv = tv; ci = ti;
		}
	}
}

// Merge the chunk lists of one micrograph into its top-k candidate indices:
// 64 threads merge interleaved chunks, thread 0 merges their 64 lists.
__global__ __launch_bounds__(64) void k_grid_merge(const float2* __restrict__ part, int nchunks, int np,
	int* __restrict__ top)
{
	__shared__ float sv[64][kGridTopK];
	__shared__ int si[64][kGridTopK];
	const int col = blockIdx.x, t = threadIdx.x;
	float v[kGridTopK];
	int ix[kGridTopK];
#pragma unroll
	for (int i = 0; i < kGridTopK; ++i) { v[i] = -3.0e38f; ix[i] = 0x7fffffff; }
	for (int c = t; c < nchunks; c += 64) {
		const float2* p = part + ((long long)c * np + col) * kGridTopK;
		for (int q = 0; q < kGridTopK; ++q) {
			const int ci = __float_as_int(p[q].y);
			if (ci == 0x7fffffff) break;
			topk_insert(v, ix, p[q].x, ci);
		}
	}
#pragma unroll
	for (int i = 0; i < kGridTopK; ++i) { sv[t][i] = v[i]; si[t][i] = ix[i]; }
	__syncthreads();
	if (t != 0) return;
	for (int u = 1; u < 64; ++u)
		for (int q = 0; q < kGridTopK; ++q) {
			if (si[u][q] == 0x7fffffff) break;
			topk_insert(v, ix, sv[u][q], si[u][q]);
		}
	for (int i = 0; i < kGridTopK; ++i) top[col * kGridTopK + i] = ix[i];
}

// FP32 re-score of the top-k candidates with the 2D-search evaluator;
// grid (kGridTopK / 8, kPixChunks, images).  part: [image][k][chunk].
__global__ __launch_bounds__(256) void k_grid_rescore(GridSpec g, CtfConsts cc, RingTables ring,
	const float* __restrict__ ring_spec, const int* __restrict__ top, float ext_phase_deg,
	double2* __restrict__ part)
{
	constexpr int CPB = 8;
	__shared__ CandCtf cand[CPB];
	__shared__ double red[2][CPB][4];
	const int img = blockIdx.z;
	if (threadIdx.x < CPB) {
		const int m = top[img * kGridTopK + blockIdx.x * CPB + threadIdx.x];
		float dfm = 0.f, ast = 0.f, az = 0.f;
		if (m < g.m) grid_candidate(g, m, dfm, ast, az);
		cand[threadIdx.x] = cand_ctf(cc, dfm, ast, az, ext_phase_deg);
	}
	__syncthreads();
	const int chunk = blockIdx.y;
	const int per = (ring.n + kPixChunks - 1) / kPixChunks;
	const int k0 = chunk * per, k1 = min(ring.n, k0 + per);
	const float* sp = ring_spec + (long long)img * ring.n;
	float acs[CPB], acc2[CPB];
	CandCtf cd[CPB];
#pragma unroll
	for (int j = 0; j < CPB; ++j) { acs[j] = 0.f; acc2[j] = 0.f; cd[j] = cand[j]; }
	for (int k = k0 + threadIdx.x; k < k1; k += 256) {
		const float a = ring.a[k], b = ring.b[k], c2 = ring.c2[k], s2 = ring.s2[k], w = ring.w[k], sv = sp[k];
#pragma unroll
		for (int j = 0; j < CPB; ++j) {
			float gg = fmaf(c2, cd[j].ca, s2 * cd[j].sa);
			float df = fmaf(cd[j].sig, gg, cd[j].mean);
			float u = fmaf(a, df, cd[j].ph0 - b);
			float cv = (abs_sinpi(u) - 0.5f) * w;
			acs[j] = fmaf(cv, sv, acs[j]);
			acc2[j] = fmaf(cv, cv, acc2[j]);
		}
	}
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
	for (int j = 0; j < CPB; ++j) {
		double x = acs[j], y = acc2[j];
		for (int o = 32; o > 0; o >>= 1) { x += __shfl_xor(x, o); y += __shfl_xor(y, o); }
		if (lane == 0) { red[0][j][wave] = x; red[1][j][wave] = y; }
	}
	__syncthreads();
	if (threadIdx.x < CPB) {
		const int j = threadIdx.x, c = blockIdx.x * CPB + j;
		double x = red[0][j][0] + red[0][j][1] + red[0][j][2] + red[0][j][3];
		double y = red[1][j][0] + red[1][j][1] + red[1][j][2] + red[1][j][3];
		part[((long long)img * kGridTopK + c) * kPixChunks + chunk] = make_double2(x, y);
	}
}

// Best re-scored candidate per micrograph -> SearchState results.
__global__ void k_grid_pick(GridSpec g, const double2* __restrict__ part, const double* __restrict__ ring_ss,
	const int* __restrict__ top, float ext_phase_deg, SearchState* __restrict__ st, float* __restrict__ rescored,
	int nimg)
{
	const int img = blockIdx.x * blockDim.x + threadIdx.x;
	if (img >= nimg) return;
	float best = (float)-1e20;
	int bm = -1;
	for (int c = 0; c < kGridTopK; ++c) {
		const int m = top[img * kGridTopK + c];
		double x = 0.0, y = 0.0;
		for (int k = 0; k < kPixChunks; ++k) {
			double2 v = part[((long long)img * kGridTopK + c) * kPixChunks + k];
			x += v.x; y += v.y;
		}
		const double d = sqrt(y * ring_ss[img]);
		const float ccv = d == 0.0 ? 0.0f : (float)(x / d);
		if (rescored) rescored[img * kGridTopK + c] = ccv;
		if (m >= g.m) continue;
		if (ccv > best || (ccv == best && m < bm)) { best = ccv; bm = m; }
	}
	SearchState s = st[img];
	if (bm >= 0) {
		float dfm, ast, az;
		grid_candidate(g, bm, dfm, ast, az);
		s.dfmean = dfm; s.ast = ast; s.ang = az; s.phase = ext_phase_deg;
		s.df_min = dfm * (1.0f - ast);
		s.df_max = dfm * (1.0f + ast);
		s.ast_ang = az;
		s.ext_phase = ext_phase_deg;
		s.score = best;
		s.ccmax = best;
	}
	st[img] = s;
}

}  // namespace gctf
