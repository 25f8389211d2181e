// k_search.hpp -- correlation-driven defocus search on the device.
//
// The candidate sequences, step counts, ranges and acceptance rules are the
// reference's (CFindDefocus1D.cpp, CFindCtf1D.cpp, CFindDefocus2D.cpp),
// evaluated in IEEE single with no contraction (this library is built with
// -ffp-contract=off), so every candidate value is bit-identical to a host
// evaluation of the same float expressions.  The correlation sums run in
// FP32 per thread and FP64 across threads and pixel chunks.
#pragma once
#include "gctf_internal.hpp"

namespace gctf {

constexpr float kD2R2D = 0.01745329f;    // CFindDefocus2D.cpp:11
constexpr float kD2R1D = 0.0174533f;     // CFindDefocus1D.cpp:11
constexpr float kPi1D = 3.141592654f;    // GCalcCTF1D.cu:30
constexpr float kPi2D = 3.1415926f;      // GCalcCTF2D.cu:35

// |sin(pi u)| - 1/2 from v = u - 1/2: |sin(pi u)| = cos(pi r) with
// r = v - rint(v) in [-1/2, 1/2] (exact), and cos(pi r) - 1/2 = p(r^2) with p
// of degree 5: a minimax fit of cos(pi sqrt z) - 1/2 on z in [0, 1/4]
// (Lawson iteration on 4000 Chebyshev nodes; |error| < 2.2e-10).  In FP32
// with fused multiply-adds the result is within 1.5e-7 of the exact value,
// as close as the degree-6 Taylor series, with one FMA less.
__device__ __forceinline__ float sinpi_dev(float v)
{
	const float r = v - rintf(v);
	const float z = r * r;
	float p = fmaf(z, -2.4396716595e-02f, 2.3493755895e-01f);
	p = fmaf(z, p, -1.3352120567e+00f);
	p = fmaf(z, p, 4.0587091643e+00f);
	p = fmaf(z, p, -4.9348021372e+00f);
	return fmaf(z, p, 0.5f);
}

// ---------------------------------------------------------------------------
// 1D search (CFindDefocus1D::mBrutalForceSearch, CFindDefocus1D.cpp:76-119;
// CFindCtf1D::mFindDefocus / mRefineDefocus, CFindCtf1D.cpp:73-111).
// Each brute-force stage is one k_s1d_eval launch over all candidates of all
// micrographs (a wavefront per candidate group, lanes over the radial bins),
// then one k_s1d_pick per micrograph takes the first strict maximum and sets
// up the next stage.
// ---------------------------------------------------------------------------
struct Search1DParams {
	int cmpx;
	int mode;                // 0: Do1D, 1: Refine1D(init_df, range)
	float init_df, range;    // Refine1D arguments
	// initial Do1D scan (mFindDefocus): defocus [scan_lo, scan_hi] A in
	// steps of scan_step A; 0 keeps the reference's 3000 p^2 .. 30000 p^2
	// in 500 steps (CFindCtf1D.cpp:79-80, CFindDefocus1D.cpp:76-80)
	float scan_lo, scan_hi, scan_step;
};

// Candidates per wavefront of k_s1d_eval; 4 wavefronts per workgroup.
constexpr int kS1dCand = 4;
constexpr int kS1dPerWg = 4 * kS1dCand;

// The brute-force lattice of CFindDefocus1D::mBrutalForceSearch
// (CFindDefocus1D.cpp:76-96): defocus [dfr0, dfr1] in 500 steps of at least
// 10 A (or `step` when > 0, the configured scan), phase [phr0, phr1] in 36
// steps; the point counts follow (int)(range / step) / 2 * 2 + 1.
__device__ __forceinline__ void s1d_lattice(SearchState& S, float dfr0, float dfr1, float phr0, float phr1,
	float step)
{
	int ndf = 501;
	const float dfrange = dfr1 - dfr0;
	float dfstep = dfrange / (float)(ndf - 1);
	if (dfstep < 10) dfstep = 10.0f;
	if (step > 0.0f) dfstep = step;
	ndf = (int)(dfrange / dfstep) / 2 * 2 + 1;
	int nps = 37;
	const float psrange = phr1 - phr0;
	const float psstep = psrange / (float)(nps - 1);
	if (psstep == 0) nps = 1;
	else nps = (int)(psrange / psstep) / 2 * 2 + 1;
	S.s1_df0 = dfr0; S.s1_dfstep = dfstep; S.s1_ndf = ndf;
	S.s1_ph0 = phr0; S.s1_phstep = psstep;
	S.s1_n = ndf * nps;
}

// Candidate i's defocus (A) and phase (degrees, clamped to [0, 150]).
__device__ __forceinline__ void s1d_candidate(const SearchState& S, int i, float& df, float& ph)
{
	const int ifo = i % S.s1_ndf, iph = i / S.s1_ndf;
	df = S.s1_df0 + (float)ifo * S.s1_dfstep;
	ph = S.s1_ph0 + (float)iph * S.s1_phstep;
	if (ph < 0) ph = 0.0f;
	else if (ph > 150.0f) ph = 150.0f;
}

// CFindCtf1D::mRefineDefocus (CFindCtf1D.cpp:93-111) around df_min with the
// phase window built from the zero-initialised slot (CFindCtf1D.cpp:104).
__device__ __forceinline__ void s1d_refine_lattice(SearchState& S, float df_min, float ext_phase, float range,
	float pix)
{
	const float mindf = 3000.0f * pix * pix;
	const float dfr0 = fmaxf(df_min - range / 2, mindf);
	const float dfr1 = dfr0 + range;
	const float prange = 0.2f * S.phase_range;
	const float phr0 = fmaxf(ext_phase - prange / 2, 0.0f);
	const float phr1 = fminf(0.0f + prange, 150.0f);
	s1d_lattice(S, dfr0, dfr1, phr0, phr1, 0.0f);
	S.s1_stage = 1;
}

// First stage of Do1D (mFindDefocus) or Refine1D (mRefineDefocus directly).
__global__ void k_s1d_begin(Search1DParams p, const ImgPlan* __restrict__ plans, SearchState* __restrict__ st,
	int n)
{
	const int b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b >= n) return;
	SearchState& S = st[b];
	if (S.skip) return;
	const float pix = plans[b].cc.pix;
	if (p.mode == 0) {
		const float p2 = pix * pix;
		const float dfr0 = p.scan_lo > 0.0f ? p.scan_lo : 3000.0f * p2;
		const float dfr1 = p.scan_hi > 0.0f ? p.scan_hi : 30000.0f * p2;
		const float phr0 = fmaxf(S.ext_phase - S.phase_range / 2, 0.0f);
		const float phr1 = fminf(phr0 + S.phase_range, 150.0f);
		s1d_lattice(S, dfr0, dfr1, phr0, phr1, p.scan_step);
		S.s1_stage = 0;
	} else {
		s1d_refine_lattice(S, p.init_df, S.ext_phase, p.range, pix);
	}
}

// CC of the pending stage's candidates (GCalcCTF1D.cu:23-32, GCC1D.cu:15-125):
// per radial bin x in [lo, hi), c = (|sin(chi)| - 1/2) exp(-(x/256)^2) with
// chi / pi = ph0 + a_x df - b_x in IEEE single (a_x = (pi'/pi) wl f^2,
// b_x = (pi'/pi) wl^3 f^4 Cs / 2, pi' the kernel's 3.141592654f), the sums
// c S and c^2 in FP32 per lane and FP64 across lanes, CC = sum cS /
// sqrt(sum c^2 sum S^2) (0 when a sum is <= 0).  Each workgroup keeps the
// first strict maximum of its kS1dPerWg consecutive candidates:
// part[image][workgroup] = (CC, candidate index as int bits).  Grid
// (ceil(max candidates / kS1dPerWg), images), 256 threads.
__global__ __launch_bounds__(256) void k_s1d_eval(int cmpx, const ImgPlan* __restrict__ plans,
	const float* __restrict__ radial, const double* __restrict__ w1d, const SearchState* __restrict__ st,
	int ldp, float2* __restrict__ part)
{
	__shared__ float s_cc[kS1dPerWg];
	const SearchState& S = st[blockIdx.y];
	if (S.skip || S.s1_stage > 1) return;
	const int n = S.s1_n;
	if ((int)blockIdx.x * kS1dPerWg >= n) return;
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	const int c0 = blockIdx.x * kS1dPerWg + wave * kS1dCand;
	const ImgPlan P = plans[blockIdx.y];
	const float* rad = radial + (long long)blockIdx.y * cmpx;
	const double wl = P.cc.wl, cs = P.cc.cs;
	const double pir = (double)kPi1D / 3.14159265358979323846;
	float dfp[kS1dCand], ph0h[kS1dCand];
#pragma unroll
	for (int j = 0; j < kS1dCand; ++j) {
		float df, ph;
		s1d_candidate(S, min(c0 + j, n - 1), df, ph);
		dfp[j] = df / P.cc.pix;                                    // CFindDefocus1D.cpp:125
		const float add = P.cc.amp_phase + ph * kD2R1D;             // :124, GCalcCTF1D.cu:63
		ph0h[j] = (float)((double)add / 3.14159265358979323846 - 0.5);
	}
	float acs[kS1dCand], acc[kS1dCand], ass = 0.f;
#pragma unroll
	for (int j = 0; j < kS1dCand; ++j) { acs[j] = 0.f; acc[j] = 0.f; }
	for (int x = lane; x < cmpx; x += 64) {
		if (!((float)x >= P.lo1d && (float)x < P.hi1d)) continue;
		double f2 = (double)x * 0.5 / (double)(cmpx - 1);
		f2 *= f2;
		const float a = (float)(pir * wl * f2);
		const float b = (float)(pir * 0.5 * wl * wl * wl * f2 * f2 * cs);
		const float w = (float)w1d[x];
		const float sv = rad[x];
#pragma unroll
		for (int j = 0; j < kS1dCand; ++j) {
			const float cv = sinpi_dev(fmaf(a, dfp[j], ph0h[j] - b)) * w;
			acs[j] = fmaf(cv, sv, acs[j]);
			acc[j] = fmaf(cv, cv, acc[j]);
		}
		ass = fmaf(sv, sv, ass);
	}
	double dss = ass;
	for (int o = 32; o > 0; o >>= 1) dss += __shfl_xor(dss, o);
#pragma unroll
	for (int j = 0; j < kS1dCand; ++j) {
		double dcs = acs[j], dcc = acc[j];
		for (int o = 32; o > 0; o >>= 1) {
			dcs += __shfl_xor(dcs, o);
			dcc += __shfl_xor(dcc, o);
		}
		if (lane == 0) s_cc[wave * kS1dCand + j] = (dcc > 0 && dss > 0) ? (float)(dcs / sqrt(dcc * dss)) : 0.0f;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		float best = (float)-1e20;
		int bi = -1;
		for (int k = 0; k < kS1dPerWg; ++k) {
			const int idx = blockIdx.x * kS1dPerWg + k;
			if (idx < n && s_cc[k] > best) { best = s_cc[k]; bi = idx; }
		}
		part[(long long)blockIdx.y * ldp + blockIdx.x] = make_float2(best, __int_as_float(bi));
	}
}

// First strict maximum of the stage (CFindDefocus1D.cpp:97-111: phase
// outer, defocus inner, `>`) from the workgroups' maxima, then the next
// stage: mFindDefocus hands its defocus and phase to mRefineDefocus
// (CFindCtf1D.cpp:73-92); the refine stage's maximum is the result (dfMin =
// dfMax, score = its CC).  One workgroup of 256 per micrograph.
__global__ __launch_bounds__(256) void k_s1d_pick(SearchState* __restrict__ st, const ImgPlan* __restrict__ plans,
	int ldp, const float2* __restrict__ part)
{
	__shared__ float s_cc[256];
	__shared__ int s_i[256];
	SearchState& S = st[blockIdx.x];
	if (S.skip || S.s1_stage > 1) return;
	const float2* pp = part + (long long)blockIdx.x * ldp;
	const int nwg = (S.s1_n + kS1dPerWg - 1) / kS1dPerWg;
	float best = (float)-1e20;
	int bi = -1;
	for (int k = threadIdx.x; k < nwg; k += 256) {
		const float2 v = pp[k];
		const int i = __float_as_int(v.y);
		if (i >= 0 && v.x > best) { best = v.x; bi = i; }
	}
	s_cc[threadIdx.x] = best;
	s_i[threadIdx.x] = bi;
	__syncthreads();
	for (int o = 128; o > 0; o >>= 1) {
		if (threadIdx.x < o) {
			const float c2 = s_cc[threadIdx.x + o], c1 = s_cc[threadIdx.x];
			const int i2 = s_i[threadIdx.x + o], i1 = s_i[threadIdx.x];
			if (i2 >= 0 && (i1 < 0 || c2 > c1 || (c2 == c1 && i2 < i1))) {
				s_cc[threadIdx.x] = c2; s_i[threadIdx.x] = i2;
			}
		}
		__syncthreads();
	}
	if (threadIdx.x != 0) return;
	const int i = s_i[0];
	float df = 0.f, ph = 0.f, score = (float)-1e20;
	if (i >= 0) {
		s1d_candidate(S, i, df, ph);
		score = s_cc[0];
	}
	if (S.s1_stage == 0) {
		// CFindCtf1D::Do1D: refine over max(0.3 df, 3000) A around the scan's best
		s1d_refine_lattice(S, df, ph, fmaxf(0.3f * df, 3000.0f), plans[blockIdx.x].cc.pix);
		return;
	}
	S.ext_phase = ph;
	S.df_min = df;
	S.df_max = df;
	S.score = score;
	S.s1_stage = 2;
}

// ---------------------------------------------------------------------------
// 2D search (CFindDefocus2D.cpp:100-380).
// ---------------------------------------------------------------------------

// Candidate c of the pending stage -> (dfmean, ast, ang, phase); false for
// phase candidates outside the window (CFindDefocus2D.cpp:347).
__device__ __forceinline__ bool stage_candidate(const SearchState& s, int c, float& dfm, float& ast,
	float& ang, float& ph)
{
	dfm = s.dfmean; ast = s.ast; ang = s.ang; ph = s.phase;
	switch (s.st_kind) {
	case ST_ASTIG: {
		int j = c / s.st_nast, i = c % s.st_nast;
		ang = s.st_b0 + (float)j * s.st_bstep;
		ast = s.st_a0 + (float)i * s.st_astep;
		return true;
	}
	case ST_ASTMAG: ast = s.st_a0 + (float)c * s.st_astep; return true;
	case ST_ASTANG: ang = s.st_a0 + (float)c * s.st_astep; return true;
	case ST_DFMEAN: dfm = s.st_a0 + (float)c * s.st_astep; return true;
	case ST_PHASE:
		ph = s.phase + (float)(c - s.st_n / 2) * s.st_astep;
		return !(ph < s.st_lo || ph > s.st_hi);
	default: return true;
	}
}

// Per-candidate CTF parameters (CFindDefocus2D::mCorrelate :363-380 and
// GCalcCTF2D::DoIt :99-114) in the form the evaluator uses:
// u = chi / pi = ph0 + a * (mean + sig * g) - b with g = cos(2 (theta - az));
// ph0h = ph0 - 1/2, the offset of v = u - 1/2 that sinpi_dev takes.
struct CandCtf {
	float mean, sig, ca, sa, ph0, ph0h;
};

__device__ __forceinline__ CandCtf cand_ctf(const CtfConsts& c, float dfm, float ast, float ang, float ph)
{
	float ph_rad = ph * kD2R2D;
	float az = ang * kD2R2D;
	float dmin = dfm * (1.0f - ast) / c.pix;
	float dmax = dfm * (1.0f + ast) / c.pix;
	CandCtf o;
	o.mean = 0.5f * (dmin + dmax);
	o.sig = 0.5f * (dmax - dmin);
	float add = c.amp_phase + ph_rad;
	double a2 = 2.0 * (double)az;
	o.ca = (float)cos(a2);
	o.sa = (float)sin(a2);
	o.ph0 = (float)((double)add / 3.14159265358979323846);
	o.ph0h = (float)((double)add / 3.14159265358979323846 - 0.5);
	return o;
}

// Contribution of one ring pixel to CPB candidates' sums (c*S, c*c), with
// c = (|sin(pi u)| - 1/2) w.  SHARED: the candidates have one azimuth and
// one phase, so g and ph0h - b are formed once per pixel; the values are
// those of the general form.
template <int CPB, bool SHARED>
__device__ __forceinline__ void cc2d_pixel(const CandCtf* cd, float a, float b, float c2, float s2,
	float w, float sv, float* acc_cs, float* acc_cc)
{
	if constexpr (SHARED) {
		const float g = fmaf(c2, cd[0].ca, s2 * cd[0].sa);
		const float pb = cd[0].ph0h - b;
#pragma unroll
		for (int j = 0; j < CPB; ++j) {
			const float df = fmaf(cd[j].sig, g, cd[j].mean);
			const float cv = sinpi_dev(fmaf(a, df, pb)) * w;
			acc_cs[j] = fmaf(cv, sv, acc_cs[j]);
			acc_cc[j] = fmaf(cv, cv, acc_cc[j]);
		}
	} else {
#pragma unroll
		for (int j = 0; j < CPB; ++j) {
			const float g = fmaf(c2, cd[j].ca, s2 * cd[j].sa);
			const float df = fmaf(cd[j].sig, g, cd[j].mean);
			const float cv = sinpi_dev(fmaf(a, df, cd[j].ph0h - b)) * w;
			acc_cs[j] = fmaf(cv, sv, acc_cs[j]);
			acc_cc[j] = fmaf(cv, cv, acc_cc[j]);
		}
	}
}

// A pointer read from memory (here: the ring tables of a plan) is generic to
// the compiler, whose loads through it are then flat loads, with 64-bit
// address arithmetic and waits on the LDS counter too.  The tables are
// device memory: gload reads them as global memory, from a uniform base and
// a 32-bit index.
typedef const float __attribute__((address_space(1)))* gfloat_ptr;
__device__ __forceinline__ float gload(const float* p, int i) { return ((gfloat_ptr)p)[i]; }

// The pixel loop is unrolled so that the loads of four pixels are in flight
// together: the small stages run few waves per SIMD, which cannot hide the
// latency of one pixel's loads behind another wave's arithmetic.
template <int CPB, bool SHARED>
__device__ __forceinline__ void cc2d_pixels(const CandCtf* cd, const RingTables& ring, const float* sp,
	int k0, int k1, float* acc_cs, float* acc_cc)
{
#pragma unroll 4
	for (int k = k0 + (int)threadIdx.x; k < k1; k += 256)
		cc2d_pixel<CPB, SHARED>(cd, gload(ring.a, k), gload(ring.b, k), gload(ring.c2, k), gload(ring.s2, k),
			gload(ring.w, k), gload(sp, k), acc_cs, acc_cc);
}

// Candidate evaluation.  Grid (ceil(max candidates / CPB), kPixChunks,
// images); 256 threads.  part[image][cand][chunk] = {sum c*S, sum c*c}.
template <int CPB>
__global__ __launch_bounds__(256) void k_cc2d_eval(const ImgPlan* __restrict__ plans,
	const float* __restrict__ ring_spec, int kstride, const SearchState* __restrict__ st,
	double2* __restrict__ part)
{
	const CtfConsts cc = plans[blockIdx.z].cc;
	const RingTables ring = plans[blockIdx.z].ring;
	__shared__ CandCtf cand[CPB];
	__shared__ int cvalid[CPB];
	__shared__ double red[2][CPB][4];
	const SearchState& S = st[blockIdx.z];
	if (S.skip || S.done || S.st_kind == ST_NONE) return;
	const int n = S.st_n;
	const int c0 = blockIdx.x * CPB;
	if (c0 >= n) return;
	if (threadIdx.x < CPB) {
		int c = c0 + threadIdx.x;
		float dfm, ast, ang, ph;
		bool ok = c < n && stage_candidate(S, c, dfm, ast, ang, ph);
		cvalid[threadIdx.x] = ok;
		cand[threadIdx.x] = ok ? cand_ctf(cc, dfm, ast, ang, ph) : CandCtf{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
	}
	__syncthreads();
	const int chunk = blockIdx.y;
	const int per = (ring.n + kPixChunks - 1) / kPixChunks;
	const int k0 = chunk * per, k1 = min(ring.n, k0 + per);
	const float* sp = ring_spec + (long long)blockIdx.z * kstride;
	float acc_cs[CPB], acc_cc[CPB];
#pragma unroll
	for (int j = 0; j < CPB; ++j) { acc_cs[j] = 0.f; acc_cc[j] = 0.f; }
	CandCtf cd[CPB];
	bool shared = true;
#pragma unroll
	for (int j = 0; j < CPB; ++j) {
		cd[j] = cand[j];
		// invalid candidates are zero and their sums are not stored
		if (c0 + j < n && cvalid[j])
			shared = shared && cd[j].ca == cand[0].ca && cd[j].sa == cand[0].sa && cd[j].ph0h == cand[0].ph0h;
	}
	if (shared && cvalid[0]) cc2d_pixels<CPB, true>(cd, ring, sp, k0, k1, acc_cs, acc_cc);
	else cc2d_pixels<CPB, false>(cd, ring, sp, k0, k1, acc_cs, acc_cc);
	// reduce: wave shuffles in double, then across the 4 waves in LDS
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
	for (int j = 0; j < CPB; ++j) {
		double x = acc_cs[j], y = acc_cc[j];
		for (int o = 32; o > 0; o >>= 1) {
			x += __shfl_xor(x, o);
			y += __shfl_xor(y, o);
		}
		if (lane == 0) { red[0][j][wave] = x; red[1][j][wave] = y; }
	}
	__syncthreads();
	if (threadIdx.x < CPB) {
		const int j = threadIdx.x, c = c0 + j;
		if (c < n) {
			double x = red[0][j][0] + red[0][j][1] + red[0][j][2] + red[0][j][3];
			double y = red[1][j][0] + red[1][j][1] + red[1][j][2] + red[1][j][3];
			part[((long long)blockIdx.z * kMaxCands + c) * kPixChunks + chunk] = make_double2(x, y);
		}
	}
}

// Gather the ring pixels of each spectrum and their FP64 energy: ring
// spectrum of image i at ring_spec + i * kstride.
__device__ __forceinline__ void ring_gather(const float* __restrict__ in, const RingTables& ring,
	float* __restrict__ out, double* __restrict__ ss, double* red)
{
	double acc = 0.0;
	for (int k = threadIdx.x; k < ring.n; k += 1024) {
		float v = in[ring.idx[k]];
		out[k] = v;
		acc += (double)v * v;
	}
	red[threadIdx.x] = acc;
	__syncthreads();
	for (int s = 512; s > 0; s >>= 1) {
		if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
		__syncthreads();
	}
	if (threadIdx.x == 0) *ss = red[0];
}

__global__ __launch_bounds__(1024) void k_ring_gather(const float* __restrict__ spect, int npix,
	const ImgPlan* __restrict__ plans, int kstride, float* __restrict__ ring_spec, double* __restrict__ ring_ss)
{
	__shared__ double red[1024];
	const RingTables ring = plans[blockIdx.x].ring;
	ring_gather(spect + (long long)blockIdx.x * npix, ring, ring_spec + (long long)blockIdx.x * kstride,
		ring_ss + blockIdx.x, red);
}

// The same with one ring for the whole batch (the grid search).
__global__ __launch_bounds__(1024) void k_ring_gather_uniform(const float* __restrict__ spect, int npix,
	RingTables ring, float* __restrict__ ring_spec, double* __restrict__ ring_ss)
{
	__shared__ double red[1024];
	ring_gather(spect + (long long)blockIdx.x * npix, ring, ring_spec + (long long)blockIdx.x * ring.n,
		ring_ss + blockIdx.x, red);
}

// ---- controller -----------------------------------------------------------
// Stage preparation, exactly the float code of the reference's methods.

__device__ __forceinline__ void prep_none(SearchState& s) { s.st_kind = ST_NONE; s.st_n = 0; }

// mFindAstig (CFindDefocus2D.cpp:185-200)
__device__ void prep_astig(SearchState& s)
{
	const float tiny = (float)1e-20;
	float astr = s.astrange[1] - s.astrange[0];
	float angr = s.angrange[1] - s.angrange[0];
	if (astr < tiny && angr < tiny) { prep_none(s); return; }
	int nast = s.pg.n_ast, nang = s.pg.n_ang;
	float aststep = astr / (float)nast;
	float angstep = angr / (float)nang;
	if ((double)aststep < 1e-5) nast = 1;
	if ((double)angstep < 1e-5) nang = 1;
	if (nast == 1 && nang == 1) { prep_none(s); return; }
	s.st_kind = ST_ASTIG;
	s.st_n = nast * nang;
	s.st_nast = nast;
	s.st_a0 = s.astrange[0]; s.st_astep = aststep;
	s.st_b0 = s.angrange[0]; s.st_bstep = angstep;
}

// mRefineAstMag / mRefineAstAng (CFindDefocus2D.cpp:224-286)
__device__ void prep_line(SearchState& s, int kind, float cur, const float* lim, float range,
	int n0, float minstep)
{
	const float tiny = (float)1e-20;
	if (range < tiny) { prep_none(s); return; }
	float mn = fmaxf(cur - range * 0.5f, lim[0]);
	float mx = fminf(cur + range * 0.5f, lim[1]);
	int n = n0;
	float step = (mx - mn) / (float)(n - 1);
	if (step < minstep) step = minstep;
	n = (int)((mx - mn) / step) / 2 * 2 + 1;
	if (n <= 1) { prep_none(s); return; }
	s.st_kind = kind; s.st_n = n; s.st_a0 = mn; s.st_astep = step;
}

// mRefineDfMean (CFindDefocus2D.cpp:288-322)
__device__ void prep_dfmean(SearchState& s, float range)
{
	const float tiny = (float)1e-30;
	if (range < tiny) { prep_none(s); return; }
	int n = s.pg.n_df;
	float mn = fmaxf(s.dfmean - range * 0.5f, s.dfrange[0]);
	float mx = fminf(s.dfmean + range * 0.5f, s.dfrange[1]);
	float step = (mx - mn) / (float)n;
	if (step < s.pg.df_minstep) step = s.pg.df_minstep;
	n = (int)((mx - mn) / step) / 2 * 2 + 1;
	if (n <= 1) { prep_none(s); return; }
	s.st_kind = ST_DFMEAN; s.st_n = n; s.st_a0 = mn; s.st_astep = step;
	s.st_old = s.dfmean;
}

// mRefinePhase (CFindDefocus2D.cpp:324-361)
__device__ void prep_phase(SearchState& s, float range)
{
	const float tiny = (float)1e-30;
	if (range < tiny) { prep_none(s); return; }
	int n = 61;
	float step = range / (float)n;
	if (step < 0.03f) { prep_none(s); return; }
	if (step > 1.0f) {
		step = 1.0f;
		n = (int)(range / step);
		n = n / 2 * 2 + 1;
	}
	float mnp = s.phrange[0] - s.phrange[1] * 0.5f;
	float mxp = s.phrange[0] + s.phrange[1] * 0.5f;
	if (mnp < 0) mnp = 0.0f;
	if (mxp > 150) mxp = 150.0f;
	if (n > kMaxCands) n = kMaxCands;
	s.st_kind = ST_PHASE; s.st_n = n; s.st_astep = step; s.st_lo = mnp; s.st_hi = mxp;
}

__device__ __forceinline__ float iter_scale(const SearchState& s, int i)
{
	return 1.0f - (float)i * 0.5f / (float)s.pg.iters;   // CFindDefocus2D.cpp:167
}

// Learned path's last stage: s.pg.fine_n defocus candidates s.pg.fine_step
// apart around the current mean defocus, inside the window.
__device__ void prep_fine(SearchState& s)
{
	if (s.pg.fine_n < 2) { prep_none(s); return; }
	const int h = s.pg.fine_n / 2;
	float mn = fmaxf(s.dfmean - (float)h * s.pg.fine_step, s.dfrange[0]);
	float mx = fminf(s.dfmean + (float)h * s.pg.fine_step, s.dfrange[1]);
	int n = (int)((mx - mn) / s.pg.fine_step) + 1;
	if (n <= 1) { prep_none(s); return; }
	s.st_kind = ST_DFMEAN; s.st_n = n; s.st_a0 = mn; s.st_astep = s.pg.fine_step;
	s.st_old = s.dfmean;
}

// Stage program of mIterate; `next` names what the controller prepares.
enum NextStage : int {
	NX_BEGIN_DO2D = 0,   // CFindDefocus2D::DoIt setup, then INIT
	NX_BEGIN_REFINE = 1, // CFindDefocus2D::Refine setup, then INIT
	NX_ASTIG = 2,        // consume INIT, prepare ASTIG
	NX_ASTMAG = 3,       // consume ASTIG / PHASE, prepare ASTMAG(iter)
	NX_ASTANG = 4,       // consume ASTMAG, prepare ASTANG
	NX_DFMEAN = 5,       // consume ASTANG, prepare DFMEAN
	NX_PHASE = 6,        // consume DFMEAN, prepare PHASE
	NX_FINISH = 7,       // consume PHASE of the last iteration
	NX_BEGIN_PROPOSAL = 8, // learned path: window around the proposal, DoIt caps, then INIT
	NX_FINE = 9,         // consume PHASE of the last iteration, prepare the fine defocus stage
};

struct RefineArgs {
	float dfmean[2], ast[2], ang[2], phase[2];
};

// Prepare stage `next` of state s (prev_kind: the kind of the stage just
// consumed) and, at the end of the search, write the results.
__device__ void advance_state(SearchState& s, int prev_kind, int next)
{
	if (!s.done) {
		switch (next) {
		case NX_ASTIG:
			prep_astig(s);
			break;
		case NX_ASTMAG:
			if (prev_kind == ST_ASTIG || s.iter == 0) {
				// after mFindAstig (CFindDefocus2D.cpp:158-166)
				if (s.oldcc > s.ccmax) { s.done = 1; break; }
				s.phase = s.phrange[0];
				s.tot_df = s.dfrange[1] - s.dfrange[0];
				s.tot_ast = s.astrange[1] - s.astrange[0];
				s.tot_ang = s.angrange[1] - s.angrange[0];
			}
			s.iter += 1;
			prep_line(s, ST_ASTMAG, s.ast, s.astrange, iter_scale(s, s.iter) * s.tot_ast, s.pg.n_mag, 0.001f);
			break;
		case NX_ASTANG:
			// CFindDefocus2D.cpp:171-173
			if (s.oldcc > s.ccmax) { s.done = 1; break; }
			s.ccmax = s.oldcc;
			prep_line(s, ST_ASTANG, s.ang, s.angrange, iter_scale(s, s.iter) * s.tot_ang, s.pg.n_angl, 1.0f);
			break;
		case NX_DFMEAN:
			prep_dfmean(s, iter_scale(s, s.iter) * s.tot_df);
			break;
		case NX_PHASE:
			prep_phase(s, iter_scale(s, s.iter) * s.phrange[1]);
			break;
		case NX_FINE:
			prep_fine(s);
			break;
		case NX_FINISH:
			prep_none(s);
			break;
		default: break;
		}
	}
	if (next == NX_FINISH || s.done) {
		// CFindDefocus2D::GetDfMin/GetDfMax/... (CFindDefocus2D.cpp:38-67)
		s.df_min = s.dfmean * (1.0f - s.ast);
		s.df_max = s.dfmean * (1.0f + s.ast);
		s.ast_ang = s.ang;
		s.ext_phase = s.phase;
		s.score = s.ccmax;
		if (s.done) prep_none(s);
	}
}

// One workgroup (256 threads) per micrograph: consume the CCs of the stage
// just evaluated (first strict maximum, `<= continue` order), update the
// state as the reference method does, and prepare the next stage (then
// next2's, when next2 >= 0 and stage `next` evaluates nothing).
__global__ __launch_bounds__(256) void k_s2d_step(SearchState* __restrict__ st,
	const double2* __restrict__ part, const double* __restrict__ ring_ss, int next, int next2,
	const RefineArgs* __restrict__ refine, SearchProgram prog)
{
	__shared__ float s_cc[256];
	__shared__ int s_i[256];
	SearchState* S = st + blockIdx.x;
	const int tid = threadIdx.x;
	if (S->skip) return;
	if (next == NX_BEGIN_DO2D || next == NX_BEGIN_REFINE || next == NX_BEGIN_PROPOSAL) {
		if (tid == 0) {
			SearchState s = *S;
			s.pg = plain_program();
			if (next == NX_BEGIN_PROPOSAL) {
				// the reference's DoIt box (CFindDefocus2D.cpp:113-118) with the
				// defocus window of the proposal
				const RefineArgs& r = refine[blockIdx.x];
				s.pg = prog;
				s.dfmean = r.dfmean[0]; s.ast = 0.0f; s.ang = 0.0f;
				s.phrange[0] = s.ext_phase;
				s.phrange[1] = s.phase_range;
				const float h = 0.5f * r.dfmean[1];
				s.dfrange[0] = fmaxf(s.dfmean - h, 3000.0f);
				s.dfrange[1] = s.dfmean + h;
				s.astrange[0] = 0.0f; s.astrange[1] = 0.06f;
				s.angrange[0] = -90.0f; s.angrange[1] = 90.0f;
			} else if (next == NX_BEGIN_DO2D) {
				// CFindCtf2D::Do2D (CFindCtf2D.cpp:44-53) + DoIt (:110-124)
				float dfmean = (s.df_min + s.df_max) * 0.5f;
				s.dfmean = dfmean; s.ast = 0.0f; s.ang = 0.0f;
				s.phrange[0] = s.ext_phase;
				s.phrange[1] = s.phase_range;
				s.dfrange[0] = s.dfmean * 0.9f; s.dfrange[1] = s.dfmean * 1.1f;
				s.astrange[0] = 0.0f; s.astrange[1] = 0.06f;
				s.angrange[0] = -90.0f; s.angrange[1] = 90.0f;
			} else {
				// CFindCtf2D::Refine (CFindCtf2D.cpp:55-68) + Refine (:126-151)
				const RefineArgs& r = refine[blockIdx.x];
				s.dfmean = r.dfmean[0]; s.ast = r.ast[0]; s.ang = r.ang[0];
				s.phrange[0] = r.phase[0];
				float h = 0.5f * r.dfmean[1];
				s.dfrange[0] = fmaxf(s.dfmean - h, 3000.0f);
				s.dfrange[1] = s.dfmean + h;
				h = 0.5f * r.ast[1];
				s.astrange[0] = fmaxf(s.ast - h, 0.0f);
				s.astrange[1] = fminf(s.ast + h, 0.02f);
				h = 0.5f * r.ang[1];
				s.angrange[0] = fmaxf(s.ang - h, -90.0f);
				s.angrange[1] = fminf(s.ang + h, 90.0f);
				s.phrange[1] = r.phase[1];
			}
			s.done = 0; s.iter = 0;
			s.st_kind = ST_INIT; s.st_n = 1;
			*S = s;
		}
		return;
	}
	// ---- argmax of the evaluated stage
	const SearchState s0 = *S;
	const bool ran = !s0.done && s0.st_kind != ST_NONE && s0.st_n > 0;
	float best = (float)-1e20;
	int bi = -1;
	if (ran) {
		const double ss = ring_ss[blockIdx.x];
		const double2* pp = part + (long long)blockIdx.x * kMaxCands * kPixChunks;
		for (int c = tid; c < s0.st_n; c += 256) {
			float dfm, ast, ang, ph;
			if (!stage_candidate(s0, c, dfm, ast, ang, ph)) continue;
			double x = 0.0, y = 0.0;
			for (int k = 0; k < kPixChunks; ++k) {
				double2 v = pp[(long long)c * kPixChunks + k];
				x += v.x; y += v.y;
			}
			double d = sqrt(y * ss);
			float cc = d == 0.0 ? 0.0f : (float)(x / d);
			if (cc > best) { best = cc; bi = c; }
		}
	}
	s_cc[tid] = best; s_i[tid] = bi;
	__syncthreads();
	for (int o = 128; o > 0; o >>= 1) {
		if (tid < o) {
			float c2 = s_cc[tid + o];
			int i2 = s_i[tid + o];
			float c1 = s_cc[tid];
			int i1 = s_i[tid];
			if (i2 >= 0 && (i1 < 0 || c2 > c1 || (c2 == c1 && i2 < i1))) {
				s_cc[tid] = c2; s_i[tid] = i2;
			}
		}
		__syncthreads();
	}
	if (tid != 0) return;
	SearchState s = s0;
	const int win = s_i[0];
	const float wcc = s_cc[0];
	if (ran && win >= 0) {
		float dfm, ast, ang, ph;
		stage_candidate(s0, win, dfm, ast, ang, ph);
		switch (s0.st_kind) {
		case ST_INIT:
			s.ccmax = wcc; s.oldcc = wcc;
			break;
		case ST_ASTIG:
			if (wcc > s.ccmax) { s.ast = ast; s.ang = ang; s.ccmax = wcc; }
			break;
		case ST_ASTMAG:
			if (wcc > s.ccmax) { s.ast = ast; s.ccmax = wcc; }
			break;
		case ST_ASTANG:
			if (wcc > s.ccmax) { s.ang = ang; s.ccmax = wcc; }
			break;
		case ST_DFMEAN:
			if (wcc > s.ccmax) { s.dfmean = dfm; s.ccmax = wcc; }
			break;
		case ST_PHASE:
			if (wcc > s.ccmax) { s.phase = ph; s.ccmax = wcc; }
			break;
		default: break;
		}
	}
	advance_state(s, s0.st_kind, next);
	// a second transition with no stage evaluated in between (the phase
	// stage of a search without phase range prepares nothing)
	if (next2 >= 0 && !s.done) advance_state(s, s.st_kind, next2);
	*S = s;
}

// ---------------------------------------------------------------------------
// The CTF at the final parameters over the half spectrum, in FP32 with the
// reference formula (GCalcCTF2D.cu:23-37, GSpectralCC2D.cu:12-109), once
// per pixel: every pixel belongs to about ten shells.  Grid (cmpy, images).
__global__ __launch_bounds__(256) void k_ctf_map(const ImgPlan* __restrict__ plans, int cmpx, int cmpy,
	const SearchState* __restrict__ st, float* __restrict__ ctfmap)
{
	const SearchState& S = st[blockIdx.y];
	if (S.skip) return;
	const CtfConsts cc = plans[blockIdx.y].cc;
	const int y = blockIdx.x;
	float ph_rad = S.phase * kD2R2D, az = S.ang * kD2R2D;
	float dmin = S.dfmean * (1.0f - S.ast) / cc.pix;
	float dmax = S.dfmean * (1.0f + S.ast) / cc.pix;
	float mean = 0.5f * (dmin + dmax), sig = 0.5f * (dmax - dmin);
	float add = cc.amp_phase + ph_rad;
	float w2 = cc.wl * cc.wl;
	float* out = ctfmap + ((long long)blockIdx.y * cmpy + y) * cmpx;
	const float fY = (float)(y - cmpy / 2) / (float)cmpy;
	for (int x = threadIdx.x; x < cmpx; x += 256) {
		float fX = (float)x * 0.5f / (float)(cmpx - 1);
		float s2 = fX * fX + fY * fY;
		float th = atanf(fY / (fX + (float)1e-30));
		float df = mean + sig * cosf(2.0f * (th - az));
		out[x] = -sinf(add + kPi2D * cc.wl * s2 * (df - 0.5f * s2 * w2 * cc.cs));
	}
}

// Spectral CC shells (GSpectralCC2D.cu:12-109): Pearson CC of CTF and
// spectrum over the pixels with lo <= |(x, y - cmpy/2)| < hi of each shell
// (width 10, clamped at the ends).  The shell bounds are integers, so a
// pixel with radius fr belongs to shell [lo, hi) exactly when
// lo <= floor(fr) < hi: the six sums of each integer radius bin are formed
// once (k_shell_bins, from a host table listing the pixels of every bin,
// radii in the kernel's float arithmetic), and each shell adds its bins.
// Grid (nbins, images), one wavefront per bin; bins[image][bin][6].
__global__ __launch_bounds__(64) void k_shell_bins(const float* __restrict__ spect,
	const float* __restrict__ ctfmap, int npix, const int* __restrict__ boff, const int* __restrict__ bcnt,
	const int* __restrict__ bidx, int nbins, const SearchState* __restrict__ st, double* __restrict__ bins)
{
	const int b = blockIdx.x, img = blockIdx.y;
	if (st[img].skip) return;
	const float* sp = spect + (long long)img * npix;
	const float* cm = ctfmap + (long long)img * npix;
	const int o = boff[b], n = bcnt[b];
	double a[6] = {0, 0, 0, 0, 0, 0};
	for (int i = threadIdx.x; i < n; i += 64) {
		const int p = bidx[o + i];
		const float c = cm[p], sv = sp[p];
		a[0] += c; a[1] += sv; a[2] += (double)c * sv; a[3] += (double)c * c; a[4] += (double)sv * sv;
	}
	a[5] = (double)n;
#pragma unroll
	for (int j = 0; j < 5; ++j)
		for (int off = 32; off > 0; off >>= 1) a[j] += __shfl_xor(a[j], off);
	if (threadIdx.x < 6) bins[((long long)img * nbins + b) * 6 + threadIdx.x] = a[threadIdx.x];
}

// Shell k = the bins [lo, hi); grid (ceil(cmpx / 64), images).
__global__ __launch_bounds__(64) void k_shell_cc(const double* __restrict__ bins, int nbins, int cmpx, int width,
	const SearchState* __restrict__ st, float* __restrict__ shell_cc)
{
	const int k = blockIdx.x * 64 + threadIdx.x, img = blockIdx.y;
	if (k >= cmpx || st[img].skip) return;
	int lo = k - width / 2, hi = lo + width;
	if (lo < 0) { lo = 0; hi = width; }
	else if (hi > cmpx) { lo = cmpx - width; hi = cmpx; }
	double a[6] = {0, 0, 0, 0, 0, 0};
	const double* bb = bins + (long long)img * nbins * 6;
	for (int r = lo; r < hi && r < nbins; ++r)
#pragma unroll
		for (int j = 0; j < 6; ++j) a[j] += bb[r * 6 + j];
	float out = 0.f;
	const double n = a[5];
	if (n > 0) {
		double mc = a[0] / n, ms = a[1] / n, mcs = a[2] / n;
		double vc = a[3] / n - mc * mc, vs = a[4] / n - ms * ms;
		if (vc > 0 && vs > 0) out = (float)((mcs - mc * ms) / sqrt(vc * vs));
	}
	shell_cc[(long long)img * cmpx + k] = out;
}

// GSpectralCC2D::DoIt shell pick (:157-171) and m_fCtfRes (CFindDefocus2D.cpp:399).
__global__ void k_ctf_res(const float* __restrict__ shell_cc, int cmpx, int cmpy,
	const ImgPlan* __restrict__ plans, SearchState* __restrict__ st, int nimg)
{
	const int b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b >= nimg || st[b].skip) return;
	const float pix = plans[b].cc.pix;
	const float* cc = shell_cc + (long long)b * cmpx;
	int imax = 1;
	float fmax = (float)-1e30;
	for (int i = 1; i < cmpx; ++i) {
		if (cc[i] <= fmax) continue;
		fmax = cc[i]; imax = i;
	}
	int shell = imax;
	if (!(fmax < 0.143f)) {
		for (int i = imax; i < cmpx; ++i) {
			if (cc[i] < 0.143f) break;
			shell = i;
		}
	}
	st[b].ctf_res = (float)cmpy * pix / (float)shell;
}

}  // namespace gctf
