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

constexpr int kGemmKs = 32;     // ring pixels per pipeline stage of the GEMM
constexpr int kGemmTabs = 8;    // per-pixel tables of a stage (6 used)
constexpr int kGemmWaves = 8;   // waves per workgroup, 32 candidates each

// Ring spectra of the batch in the B-operand layout:
// bt[(k / 16) * np + n][k % 16] as bf16, zero for k >= K or n >= nimg
// (kpad = K rounded up to kGemmKs); fn[n] = sum over k of the bf16 values
// squared (FP64).
__global__ __launch_bounds__(256) void k_grid_pack(const float* __restrict__ ring_spec, int K, int kpad,
	int nimg, int np, __bf16* __restrict__ bt, double* __restrict__ fn)
{
	__shared__ double red[256];
	const int n = blockIdx.x;
	double acc = 0.0;
	for (int k = threadIdx.x; k < kpad; k += 256) {
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

// Per-pixel tables of the GEMM's CTF generation, in revolutions of the
// hardware sine (|sin(pi u)| = |sin(2 pi (u / 2))|):
//   u / 2 = ga mean + gac (sig cos 2az) + gas (sig sin 2az) + gnb
// with ga = a / 2, gac = a cos(2 theta) / 2, gas = a sin(2 theta) / 2,
// gnb = (ph0 - b) / 2 (the grid shares one phase shift ph0), and
// c = |sin| w - w / 2.  Layout gt[stage][table][32]; zero past K.
__global__ __launch_bounds__(256) void k_grid_tables(CtfConsts cc, RingTables ring, float ext_phase_deg,
	int kpad, float* __restrict__ gt)
{
	const int k = blockIdx.x * 256 + threadIdx.x;
	if (k >= kpad) return;
	const float ph0 = cand_ctf(cc, 0.f, 0.f, 0.f, ext_phase_deg).ph0;
	float t[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
	if (k < ring.n) {
		const float a = ring.a[k];
		t[0] = 0.5f * a;
		t[1] = 0.5f * a * ring.c2[k];
		t[2] = 0.5f * a * ring.s2[k];
		t[3] = 0.5f * (ph0 - ring.b[k]);
		t[4] = ring.w[k];
		t[5] = 0.5f * ring.w[k];
	}
	float* o = gt + (long long)(k / kGemmKs) * kGemmTabs * kGemmKs + (k % kGemmKs);
#pragma unroll
	for (int i = 0; i < 6; ++i) o[i * kGemmKs] = t[i];
	o[6 * kGemmKs] = 0.f;
	o[7 * kGemmKs] = 0.f;
}

// The GEMM.  One workgroup = 8 waves = 256 candidates x NP = 32 NT
// micrograph columns; grid (ceil(M / 256)).  Every wave holds its
// 32 x NP accumulator tile for the whole K loop; the B tile of a stage
// (32 ring pixels x NP columns, bf16) and the stage's pixel tables are
// staged once per workgroup in LDS, double-buffered: the global loads of
// stage s + 1 are in flight while stage s is computed.
// cc: [M][NP] floats (m-major), CC = numerator / sqrt(E_m F_n).
template <int NT>
__global__ __launch_bounds__(512) void k_grid_gemm(GridSpec g, CtfConsts cc, const float* __restrict__ gt,
	int nst, const __bf16* __restrict__ bt, const double* __restrict__ fn, float ext_phase_deg,
	float* __restrict__ out)
{
	constexpr int NP = 32 * NT;
	constexpr int BV = NP * kGemmKs * 2 / 16;          // 16-byte vectors per B stage
	constexpr int BPT = (BV + 511) / 512;              // per thread
	constexpr int TV = kGemmTabs * kGemmKs / 4;        // float4 per table stage (64)
	__shared__ uint4 s_b[2][BV];
	__shared__ float4 s_t[2][TV];
	__shared__ float s_e[kGemmWaves][32];
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int half = lane >> 5, row = lane & 31;
	const int m = blockIdx.x * (32 * kGemmWaves) + wave * 32 + row;
	float h_mean = 0.f, h_sc = 0.f, h_ss = 0.f;
	if (m < g.m) {
		float dfm, ast, az;
		grid_candidate(g, m, dfm, ast, az);
		const CandCtf c = cand_ctf(cc, dfm, ast, az, ext_phase_deg);
		h_mean = c.mean;
		h_sc = c.sig * c.ca;
		h_ss = c.sig * c.sa;
	}
	const uint4* gb = reinterpret_cast<const uint4*>(bt);
	const float4* gtab = reinterpret_cast<const float4*>(gt);
	// this thread's share of a B stage: vectors v0 and (BV > 512 only) v1
	static_assert(BV <= 1024, "at most two B vectors per thread");
	const int v0 = tid < BV ? tid : BV - 1;
	const int v1 = tid + 512 < BV ? tid + 512 : BV - 1;
	uint4 pb0 = gb[v0], pb1 = BV > 512 ? gb[v1] : pb0;
	float4 pt = gtab[tid < TV ? tid : 0];
	if (tid < BV) s_b[0][tid] = pb0;
	if (BV > 512 && tid + 512 < BV) s_b[0][tid + 512] = pb1;
	if (tid < TV) s_t[0][tid] = pt;
	__syncthreads();
	floatx16 acc[NT], accE;
#pragma unroll
	for (int t = 0; t < NT; ++t)
#pragma unroll
		for (int r = 0; r < 16; ++r) acc[t][r] = 0.f;
#pragma unroll
	for (int r = 0; r < 16; ++r) accE[r] = 0.f;
	for (int st = 0; st < nst; ++st) {
		const int cur = st & 1;
		const bool more = st + 1 < nst;
		if (more) {
			const uint4* src = gb + (long long)(st + 1) * BV;
			pb0 = src[v0];
			if (BV > 512) pb1 = src[v1];
			pt = gtab[(long long)(st + 1) * TV + (tid < TV ? tid : 0)];
		}
		const float* tb = reinterpret_cast<const float*>(s_t[cur]);
#pragma unroll
		for (int sl = 0; sl < 2; ++sl) {
			const int k0 = sl * 16 + half * 8;
			float ga[8], gac[8], gas[8], gnb[8], w[8], hw[8];
#pragma unroll
			for (int q = 0; q < 2; ++q) {
				*reinterpret_cast<float4*>(ga + 4 * q) = *reinterpret_cast<const float4*>(tb + 0 * kGemmKs + k0 + 4 * q);
				*reinterpret_cast<float4*>(gac + 4 * q) = *reinterpret_cast<const float4*>(tb + 1 * kGemmKs + k0 + 4 * q);
				*reinterpret_cast<float4*>(gas + 4 * q) = *reinterpret_cast<const float4*>(tb + 2 * kGemmKs + k0 + 4 * q);
				*reinterpret_cast<float4*>(gnb + 4 * q) = *reinterpret_cast<const float4*>(tb + 3 * kGemmKs + k0 + 4 * q);
				*reinterpret_cast<float4*>(w + 4 * q) = *reinterpret_cast<const float4*>(tb + 4 * kGemmKs + k0 + 4 * q);
				*reinterpret_cast<float4*>(hw + 4 * q) = *reinterpret_cast<const float4*>(tb + 5 * kGemmKs + k0 + 4 * q);
			}
			bf16x8 a;
#pragma unroll
			for (int j = 0; j < 8; ++j) {
				const float u = fmaf(gac[j], h_sc, fmaf(gas[j], h_ss, fmaf(ga[j], h_mean, gnb[j])));
				a[j] = (__bf16)fmaf(fabsf(__builtin_amdgcn_sinf(u)), w[j], -hw[j]);
			}
			const char* bb = reinterpret_cast<const char*>(s_b[cur]) + sl * NP * 32 + row * 32 + half * 16;
#pragma unroll
			for (int t = 0; t < NT; ++t) {
				const bf16x8 b = *reinterpret_cast<const bf16x8*>(bb + t * 32 * 32);
				acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[t], 0, 0, 0);
			}
			// the A fragment is also the B fragment of A^T: diag(A A^T) = E_m
			accE = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, a, accE, 0, 0, 0);
		}
		if (more) {
			if (tid < BV) s_b[cur ^ 1][tid] = pb0;
			if (BV > 512 && tid + 512 < BV) s_b[cur ^ 1][tid + 512] = pb1;
			if (tid < TV) s_t[cur ^ 1][tid] = pt;
		}
		__syncthreads();
	}
	// E_m = D[c][c] of A A^T: row 8 (r / 4) + 4 h + r % 4 = c holds in lane
	// c + 32 h with h = (c % 8) / 4, register r = 4 (c / 8) + c % 4
	if (half == ((row & 7) >> 2)) {
		const int rr = 4 * (row >> 3) + (row & 3);
		float ev = 0.f;
#pragma unroll
		for (int r = 0; r < 16; ++r) ev = r == rr ? accE[r] : ev;
		s_e[wave][row] = ev;
	}
	__syncthreads();
	const int mrow0 = blockIdx.x * (32 * kGemmWaves) + wave * 32;
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
			out[(long long)mm * NP + col] = d > 0.0 ? (float)((double)acc[t][r] / d) : 0.f;
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
	// rows in groups of U: the loads of a group are independent and issue
	// back to back; the insertions then run in row order
	constexpr int U = 16;
	for (int m0 = r0; m0 < r1; m0 += U) {
		float xs[U];
#pragma unroll
		for (int u = 0; u < U; ++u) {
			const int mm = m0 + u;
			xs[u] = mm < r1 ? cc[(long long)mm * np + col] : -3.0e38f;
		}
#pragma unroll
		for (int u = 0; u < U; ++u) {
			if (!(xs[u] > v[kGridTopK - 1])) continue;
			// branch-free insertion (strictly greater values move up; equal
			// values keep their order)
			float cv = xs[u];
			int ci = m0 + u;
#pragma unroll
			for (int i = 0; i < kGridTopK; ++i) {
				const bool gt = cv > v[i];
				const float tv = v[i];
				const int ti = ix[i];
				v[i] = gt ? cv : tv;
				ix[i] = gt ? ci : ti;
				cv = gt ? tv : cv;
				ci = gt ? ti : ci;
			}
		}
	}
	float2* out = part + ((long long)blockIdx.x * np + col) * kGridTopK;
#pragma unroll
	for (int i = 0; i < kGridTopK; ++i) out[i] = make_float2(v[i], __int_as_float(ix[i]));
}

// (cc, m) total order of the candidate lists: larger cc first, then smaller m.
__device__ __forceinline__ bool topk_better(float2 x, float2 y)
{
	return x.x > y.x || (x.x == y.x && __float_as_int(x.y) < __float_as_int(y.y));
}

// dst <- the best kGridTopK of the sorted lists dst and src (both sorted
// under topk_better, padded with (-3e38, INT_MAX)).
__device__ __forceinline__ void topk_merge(float2* dst, const float2* src)
{
	float2 o[kGridTopK];
	int i = 0, j = 0;
#pragma unroll
	for (int q = 0; q < kGridTopK; ++q) {
		const float2 x = dst[i], y = src[j];
		const bool take = topk_better(x, y);
		o[q] = take ? x : y;
		i += take ? 1 : 0;
		j += take ? 0 : 1;
	}
#pragma unroll
	for (int q = 0; q < kGridTopK; ++q) dst[q] = o[q];
}

// Merge the chunk lists of one micrograph into its top-k candidate indices:
// 256 threads each fold a strided set of chunk lists into one list in LDS,
// then a tree of pairwise merges.
__global__ __launch_bounds__(256) void k_grid_merge(const float2* __restrict__ part, int nchunks, int np,
	int* __restrict__ top)
{
	__shared__ float2 sl[256][kGridTopK];
	const int col = blockIdx.x, t = threadIdx.x;
	float2* mine = sl[t];
#pragma unroll
	for (int q = 0; q < kGridTopK; ++q) mine[q] = make_float2(-3.0e38f, __int_as_float(0x7fffffff));
	for (int c = t; c < nchunks; c += 256) topk_merge(mine, part + ((long long)c * np + col) * kGridTopK);
	__syncthreads();
	for (int s = 128; s > 0; s >>= 1) {
		if (t < s) topk_merge(mine, sl[t + s]);
		__syncthreads();
	}
	if (t < kGridTopK) top[col * kGridTopK + t] = __float_as_int(sl[0][t].y);
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
			float cv = sinpi_dev(fmaf(a, df, cd[j].ph0h - b)) * w;
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
