// gctf_fft.hpp -- LDS-resident complex FFTs for gfx950.
//
// Power-of-two lengths use a Stockham autosort (radix 8, last pass radix 2
// or 4) whose butterflies live in registers; one workgroup-wide barrier
// separates the LDS reads of a pass from its writes, so one LDS buffer of N
// complex values suffices.  Lengths 2^a 3^b 5^c run the same Stockham passes
// with radix-3 and radix-5 butterflies added (fft_lds_mixed).  Other lengths
// use Bluestein's chirp-z algorithm on a power-of-two length M >= 2n-1.  All
// are unnormalised; the forward transform uses exp(-2 pi i k n / N).
//
// LDS layout: logical element i lives at physical index pidx(i) = i + i/8.
// The Stockham writes of the early passes stride by the radix (8 complex);
// the odd multiplier 9 spreads those strides over all banks.  A buffer of N
// logical elements therefore needs fft_lds_len(N) = N + N/8 entries.
//
// Twiddles: one load w = exp(-2 pi i k S / N) per butterfly from a table
// computed on the host in double; the powers w^2 .. w^(R-1) are formed by
// complex products of depth <= 3.
//
// All code here is compiled with -ffp-contract=off; fused multiply-adds
// are written out with fmaf where they are wanted.
#pragma once
#include <hip/hip_runtime.h>

namespace gctf {

__host__ __device__ constexpr int fft_lds_len(int n) { return n + n / 8; }
__device__ __forceinline__ int pidx(int i) { return i + (i >> 3); }

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
	return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}

// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b)
{
	return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -(a.x * b.y)));
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 mul_wi(float2 a)
{
	return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <int R, bool INV>
__device__ __forceinline__ void dft_small(float2* u);

template <>
__device__ __forceinline__ void dft_small<2, false>(float2* u)
{
	float2 a = u[0], b = u[1];
	u[0] = cadd(a, b); u[1] = csub(a, b);
}
template <>
__device__ __forceinline__ void dft_small<2, true>(float2* u)
{
	float2 a = u[0], b = u[1];
	u[0] = cadd(a, b); u[1] = csub(a, b);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3)
{
	float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
	float2 t2 = cadd(a1, a3), t3 = mul_wi<INV>(csub(a1, a3));
	a0 = cadd(t0, t2); a2 = csub(t0, t2);
	a1 = cadd(t1, t3); a3 = csub(t1, t3);
}

template <>
__device__ __forceinline__ void dft_small<4, false>(float2* u) { dft4<false>(u[0], u[1], u[2], u[3]); }
template <>
__device__ __forceinline__ void dft_small<4, true>(float2* u) { dft4<true>(u[0], u[1], u[2], u[3]); }

template <bool INV>
__device__ __forceinline__ void dft8(float2* u)
{
	const float c = 0.70710678118654752f;
	float2 e0 = u[0], e1 = u[2], e2 = u[4], e3 = u[6];
	float2 o0 = u[1], o1 = u[3], o2 = u[5], o3 = u[7];
	dft4<INV>(e0, e1, e2, e3);
	dft4<INV>(o0, o1, o2, o3);
	// o_k *= W8^k
	float2 w1, w3;
	if (INV) {
		w1 = make_float2(c * (o1.x - o1.y), c * (o1.x + o1.y));
		w3 = make_float2(-c * (o3.x + o3.y), c * (o3.x - o3.y));
	} else {
		w1 = make_float2(c * (o1.x + o1.y), c * (o1.y - o1.x));
		w3 = make_float2(c * (o3.y - o3.x), -c * (o3.x + o3.y));
	}
	float2 w2 = mul_wi<INV>(o2);
	u[0] = cadd(e0, o0); u[4] = csub(e0, o0);
	u[1] = cadd(e1, w1); u[5] = csub(e1, w1);
	u[2] = cadd(e2, w2); u[6] = csub(e2, w2);
	u[3] = cadd(e3, w3); u[7] = csub(e3, w3);
}

template <>
__device__ __forceinline__ void dft_small<8, false>(float2* u) { dft8<false>(u); }
template <>
__device__ __forceinline__ void dft_small<8, true>(float2* u) { dft8<true>(u); }

// u1 (+-) i s u2 parts of the radix-3 and radix-5 butterflies: y = a + mul_wi(b)
// is a - i b forward, a + i b inverse.
template <bool INV>
__device__ __forceinline__ void dft3(float2* u)
{
	const float s3 = 0.86602540378443864676f;     // sin(2 pi / 3)
	const float2 t1 = cadd(u[1], u[2]);
	const float2 t2 = make_float2(fmaf(-0.5f, t1.x, u[0].x), fmaf(-0.5f, t1.y, u[0].y));
	const float2 d = mul_wi<INV>(make_float2(s3 * (u[1].x - u[2].x), s3 * (u[1].y - u[2].y)));
	u[0] = cadd(u[0], t1);
	u[1] = cadd(t2, d);
	u[2] = csub(t2, d);
}

template <bool INV>
__device__ __forceinline__ void dft5(float2* u)
{
	const float c1 = 0.30901699437494742410f;     // cos(2 pi / 5)
	const float c2 = -0.80901699437494742410f;    // cos(4 pi / 5)
	const float s1 = 0.95105651629515357212f;     // sin(2 pi / 5)
	const float s2 = 0.58778525229247312917f;     // sin(4 pi / 5)
	const float2 a1 = cadd(u[1], u[4]), a2 = cadd(u[2], u[3]);
	const float2 b1 = csub(u[1], u[4]), b2 = csub(u[2], u[3]);
	const float2 p1 = make_float2(fmaf(c2, a2.x, fmaf(c1, a1.x, u[0].x)), fmaf(c2, a2.y, fmaf(c1, a1.y, u[0].y)));
	const float2 p2 = make_float2(fmaf(c1, a2.x, fmaf(c2, a1.x, u[0].x)), fmaf(c1, a2.y, fmaf(c2, a1.y, u[0].y)));
	const float2 q1 = mul_wi<INV>(make_float2(fmaf(s2, b2.x, s1 * b1.x), fmaf(s2, b2.y, s1 * b1.y)));
	const float2 q2 = mul_wi<INV>(make_float2(fmaf(-s1, b2.x, s2 * b1.x), fmaf(-s1, b2.y, s2 * b1.y)));
	u[0] = cadd(cadd(u[0], a1), a2);
	u[1] = cadd(p1, q1);
	u[4] = csub(p1, q1);
	u[2] = cadd(p2, q2);
	u[3] = csub(p2, q2);
}

template <>
__device__ __forceinline__ void dft_small<3, false>(float2* u) { dft3<false>(u); }
template <>
__device__ __forceinline__ void dft_small<3, true>(float2* u) { dft3<true>(u); }
template <>
__device__ __forceinline__ void dft_small<5, false>(float2* u) { dft5<false>(u); }
template <>
__device__ __forceinline__ void dft_small<5, true>(float2* u) { dft5<true>(u); }

// u[r] *= w^r for r = 1 .. R-1 (w already conjugated for the inverse).
template <int R>
__device__ __forceinline__ void twiddle_powers(float2* u, float2 w)
{
	float2 p[R];
	p[1] = w;
#pragma unroll
	for (int r = 2; r < R; ++r) {
		int hb = 1;
		while (hb * 2 <= r) hb *= 2;
		p[r] = (hb == r) ? cmul(p[r / 2], p[r / 2]) : cmul(p[hb], p[r - hb]);
	}
#pragma unroll
	for (int r = 1; r < R; ++r) u[r] = cmul(u[r], p[r]);
}

// Synchronisation of the NT threads that share one transform.  A transform
// of 64 threads belongs to a single wavefront: its LDS accesses execute in
// issue order, so a wavefront-scope fence (which keeps the compiler from
// moving LDS accesses across it and waits for the outstanding ones) replaces
// the workgroup barrier, and the other transforms of the workgroup run on.
template <int NT>
__device__ __forceinline__ void fft_sync()
{
	if constexpr (NT == 64) {
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
	} else {
		__syncthreads();
	}
}

// Twiddle sources of the power-of-two passes.  Pass P (its L > 1) of
// butterfly b of thread j uses w = exp(-2 pi i k S / N) with k = j mod L,
// S = N / (L R).  TwTable loads it from the table where the pass needs it;
// TwRegs holds every pass's values in registers, loaded once (fft_tw_load)
// by a kernel that runs many transforms of one length with the same threads:
// the passes then wait for no memory, and a load issued before the
// transform (the next transform's input) stays in flight through it.
struct TwTable {
	const float2* __restrict__ tw;
	template <int N, int R, int L, int P>
	__device__ __forceinline__ float2 get(int, int j) const { return tw[(j & (L - 1)) * (N / (L * R))]; }
};

// Passes of the power-of-two Stockham transform of length N (radix 8, the
// last 2 or 4) and the largest butterfly count per thread of any of them.
__host__ __device__ constexpr int fft_npass(int n)
{
	int p = 0;
	for (int l = 1; l < n; l *= (n / l >= 8 ? 8 : n / l)) ++p;
	return p;
}
__host__ __device__ constexpr int fft_max_bpt(int n, int nt)
{
	int m = 1;
	for (int l = 1; l < n; l *= (n / l >= 8 ? 8 : n / l)) {
		const int t = n / (n / l >= 8 ? 8 : n / l);
		m = (t + nt - 1) / nt > m ? (t + nt - 1) / nt : m;
	}
	return m;
}

template <int N, int NT>
struct TwRegs {
	float2 v[fft_npass(N)][fft_max_bpt(N, NT)];
	template <int N2, int R, int L, int P>
	__device__ __forceinline__ float2 get(int b, int) const { return v[P][b]; }
};

template <int N, int L, int NT, int P>
__device__ __forceinline__ void fft_tw_load_run(TwRegs<N, NT>& w, const float2* __restrict__ tw, int tid)
{
	if constexpr (L < N) {
		constexpr int rem = N / L;
		constexpr int R = rem >= 8 ? 8 : rem;
		if constexpr (L > 1) {
			constexpr int T = N / R;
			constexpr int BPT = (T + NT - 1) / NT;
#pragma unroll
			for (int b = 0; b < BPT; ++b) {
				const int j = tid + b * NT;
				if (T % NT == 0 || j < T) w.v[P][b] = tw[(j & (L - 1)) * (N / (L * R))];
			}
		}
		fft_tw_load_run<N, L * R, NT, P + 1>(w, tw, tid);
	}
}

// The twiddles thread tid of fft_lds<N, NT, INV> uses, from the table tw.
template <int N, int NT>
__device__ __forceinline__ void fft_tw_load(TwRegs<N, NT>& w, const float2* __restrict__ tw, int tid)
{
	fft_tw_load_run<N, 1, NT, 0>(w, tw, tid);
}

// One Stockham pass of radix R over N points; L = product of the radices of
// the previous passes; P its index.  NT threads (tid in [0, NT)) share the
// buffer.
template <int N, int R, int L, int P, int NT, bool INV, class TW>
__device__ __forceinline__ void fft_pass(float2* buf, const TW& twf, int tid)
{
	constexpr int T = N / R;
	constexpr int BPT = (T + NT - 1) / NT;
	float2 u[BPT][R];
#pragma unroll
	for (int b = 0; b < BPT; ++b) {
		const int j = tid + b * NT;
		if (T % NT == 0 || j < T) {
#pragma unroll
			for (int r = 0; r < R; ++r) u[b][r] = buf[pidx(j + r * T)];
			if constexpr (L > 1) {
				float2 w = twf.template get<N, R, L, P>(b, j);
				if (INV) w.y = -w.y;
				twiddle_powers<R>(u[b], w);
			}
			dft_small<R, INV>(u[b]);
		}
	}
	fft_sync<NT>();
#pragma unroll
	for (int b = 0; b < BPT; ++b) {
		const int j = tid + b * NT;
		if (T % NT == 0 || j < T) {
			const int k = j & (L - 1);
			const int base = (j - k) * R + k;
#pragma unroll
			for (int r = 0; r < R; ++r) buf[pidx(base + r * L)] = u[b][r];
		}
	}
	fft_sync<NT>();
}

template <int N, int L, int P, int NT, bool INV, class TW>
__device__ __forceinline__ void fft_run(float2* buf, const TW& twf, int tid)
{
	if constexpr (L < N) {
		constexpr int rem = N / L;
		constexpr int R = rem >= 8 ? 8 : rem;
		fft_pass<N, R, L, P, NT, INV>(buf, twf, tid);
		fft_run<N, L * R, P + 1, NT, INV>(buf, twf, tid);
	}
}

// In-place power-of-two FFT of the N logical elements of buf (padded
// layout, fft_lds_len(N) entries) by NT threads.  With NT == 64 the threads
// must form one wavefront and only they synchronise (fft_sync); otherwise
// every thread of the workgroup must call it.  The data must be in buf and
// visible (fft_sync<NT> or a barrier) before the call; results are visible
// to the NT threads after.
template <int N, int NT, bool INV>
__device__ __forceinline__ void fft_lds(float2* buf, const float2* __restrict__ tw, int tid)
{
	fft_run<N, 1, 0, NT, INV>(buf, TwTable{tw}, tid);
}

// The same with the twiddles already in registers (fft_tw_load).
template <int N, int NT, bool INV>
__device__ __forceinline__ void fft_lds(float2* buf, const TwRegs<N, NT>& w, int tid)
{
	fft_run<N, 1, 0, NT, INV>(buf, w, tid);
}

// ---- compile-time transform with input and output hooks -----------------
// fft_lds_io runs the same passes as fft_lds, but the first pass takes its
// inputs from ld(i) (logical input index i) instead of buf, and the last
// pass stores st(i, v) at logical output index i.  ld may read buf: every
// read of a pass precedes its writes.  Pointwise work before and after a
// transform (chirps, filters, crops) then costs no LDS round trip and no
// barrier of its own.
//
// ZIN: the inputs at i >= N/2 are zero (ld is not called for them).  DROP:
// the outputs at i >= N/2 are not needed (not stored).  In the first pass
// element r of a butterfly has index j + r N/R with j < N/R, and in the
// last pass j + r L with j < L = N/R, so both halves are r >= R/2: known
// at compile time.

// The plain input: buf itself.
struct LdsIn {
	const float2* buf;
	__device__ __forceinline__ float2 operator()(int i) const { return buf[pidx(i)]; }
};
// The plain output: v unchanged.
struct SameOut {
	__device__ __forceinline__ float2 operator()(int, float2 v) const { return v; }
};

template <int N, int R, int L, int NT, bool INV, bool ZIN, bool DROP, class Ld, class St>
__device__ __forceinline__ void fft_pass_io(float2* buf, const float2* __restrict__ tw, int tid,
	const Ld& ld, const St& st)
{
	constexpr int T = N / R;
	constexpr int BPT = (T + NT - 1) / NT;
	float2 u[BPT][R];
#pragma unroll
	for (int b = 0; b < BPT; ++b) {
		const int j = tid + b * NT;
		if (T % NT == 0 || j < T) {
#pragma unroll
			for (int r = 0; r < R; ++r) {
				if constexpr (L == 1) {
					if (ZIN && 2 * r >= R) u[b][r] = make_float2(0.f, 0.f);
					else u[b][r] = ld(j + r * T);
				} else {
					u[b][r] = buf[pidx(j + r * T)];
				}
			}
			if constexpr (L > 1) {
				const int k = j & (L - 1);
				constexpr int S = N / (L * R);
				float2 w = tw[k * S];
				if (INV) w.y = -w.y;
				twiddle_powers<R>(u[b], w);
			}
			dft_small<R, INV>(u[b]);
		}
	}
	fft_sync<NT>();
#pragma unroll
	for (int b = 0; b < BPT; ++b) {
		const int j = tid + b * NT;
		if (T % NT == 0 || j < T) {
			const int k = j & (L - 1);
			const int base = (j - k) * R + k;
#pragma unroll
			for (int r = 0; r < R; ++r) {
				if constexpr (L * R == N) {
					if (!(DROP && 2 * r >= R)) buf[pidx(base + r * L)] = st(base + r * L, u[b][r]);
				} else {
					buf[pidx(base + r * L)] = u[b][r];
				}
			}
		}
	}
	fft_sync<NT>();
}

template <int N, int L, int NT, bool INV, bool ZIN, bool DROP, class Ld, class St>
__device__ __forceinline__ void fft_run_io(float2* buf, const float2* __restrict__ tw, int tid,
	const Ld& ld, const St& st)
{
	if constexpr (L < N) {
		constexpr int rem = N / L;
		constexpr int R = rem >= 8 ? 8 : rem;
		fft_pass_io<N, R, L, NT, INV, ZIN, DROP>(buf, tw, tid, ld, st);
		fft_run_io<N, L * R, NT, INV, ZIN, DROP>(buf, tw, tid, ld, st);
	}
}

// Same contract as fft_lds, except that the input need not be in buf: the
// first pass reads ld(i).  N >= 16 (the first pass is not the last).
template <int N, int NT, bool INV, bool ZIN = false, bool DROP = false, class Ld, class St>
__device__ __forceinline__ void fft_lds_io(float2* buf, const float2* __restrict__ tw, int tid,
	const Ld& ld, const St& st)
{
	static_assert(N >= 16, "the first and last passes must differ");
	fft_run_io<N, 1, NT, INV, ZIN, DROP>(buf, tw, tid, ld, st);
}

// Bluestein DFT of length n on the compile-time working length M
// (M >= 2n-1), input element i < n given by src(i) (which may read buf),
// result in logical buf[0..n).  The chirp multiply and zero padding run in
// the forward transform's first pass, the filter multiply in its last pass
// and the final chirp multiply and 1/M scale in the inverse's last pass;
// the arithmetic is that of bluestein_lds.
// opaque(i) is i, but the compiler may not reuse an address or comparison
// computed from the same index passes earlier: keeping those live across a
// whole transform costs more registers than recomputing them.
__device__ __forceinline__ int opaque(int i)
{
	asm volatile("" : "+v"(i));
	return i;
}

template <int NT, int M, bool INV, class Src>
__device__ __forceinline__ void bluestein_io(float2* buf, int n, const float2* __restrict__ chirp,
	const float2* __restrict__ bhat, const float2* __restrict__ twm, int tid, const Src& src)
{
	// The hooks are branch-free (clamped index, then a select): branches
	// inside a pass split the live ranges of the addresses the passes share.
	auto ld = [&](int i) -> float2 {
		const int ic = min(i, n - 1);
		float2 w = chirp[ic];
		if (INV) w.y = -w.y;
		const float2 v = cmul(src(ic), w);
		return i < n ? v : make_float2(0.f, 0.f);
	};
	auto filt = [&](int k, float2 v) -> float2 {
		k = opaque(k);
		float2 b;
		if (INV) {
			b = bhat[(M - k) & (M - 1)];
			b.y = -b.y;
		} else {
			b = bhat[k];
		}
		return cmul(v, b);
	};
	// n <= M/2: the chirped input is zero from M/2 on, and only the first n
	// outputs of the inverse are kept
	fft_lds_io<M, NT, false, true, false>(buf, twm, tid, ld, filt);
	const float inv_m = 1.0f / (float)M;
	auto fin = [&](int k, float2 v) -> float2 {
		k = opaque(k);
		const int kc = min(k, n - 1);
		float2 w = chirp[kc];
		if (INV) w.y = -w.y;
		const float2 t = cmul(v, w);
		return k < n ? make_float2(t.x * inv_m, t.y * inv_m) : v;
	};
	// the inverse recomputes its addresses rather than keeping the forward's
	fft_lds_io<M, NT, true, false, true>(buf, twm, opaque(tid), LdsIn{buf}, fin);
}

// ---- runtime-length power-of-two FFT (lengths up to MAXN) ----------------
template <int R, int NT, int MAXN, bool INV>
__device__ __forceinline__ void fft_pass_rt(float2* buf, const float2* __restrict__ tw, int tid,
	int L, int T, int S)
{
	constexpr int BPT = (MAXN / R + NT - 1) / NT;
	float2 u[BPT][R];
#pragma unroll
	for (int b = 0; b < BPT; ++b) {
		const int j = tid + b * NT;
		if (j < T) {
#pragma unroll
			for (int r = 0; r < R; ++r) u[b][r] = buf[pidx(j + r * T)];
			if (L > 1) {
				const int k = j & (L - 1);
				float2 w = tw[k * S];
				if (INV) w.y = -w.y;
				twiddle_powers<R>(u[b], w);
			}
			dft_small<R, INV>(u[b]);
		}
	}
	__syncthreads();
#pragma unroll
	for (int b = 0; b < BPT; ++b) {
		const int j = tid + b * NT;
		if (j < T) {
			const int k = j & (L - 1);
			const int base = (j - k) * R + k;
#pragma unroll
			for (int r = 0; r < R; ++r) buf[pidx(base + r * L)] = u[b][r];
		}
	}
	__syncthreads();
}

// In-place power-of-two FFT of runtime length n <= MAXN; same contract as
// fft_lds.  The lengths MAXN and MAXN/2 run the compile-time transform (the
// same arithmetic, with its LDS addresses computed once per thread instead
// of once per pass); other lengths loop over runtime passes.
template <int NT, int MAXN, bool INV>
__device__ __forceinline__ void fft_lds_rt(float2* buf, int n, const float2* __restrict__ tw, int tid)
{
	if (n == MAXN) {
		fft_lds<MAXN, NT, INV>(buf, tw, tid);
		return;
	}
	if (n == MAXN / 2) {
		fft_lds<MAXN / 2, NT, INV>(buf, tw, tid);
		return;
	}
	int L = 1;
	while (L < n) {
		const int rem = n / L;
		if (rem >= 8) {
			fft_pass_rt<8, NT, MAXN, INV>(buf, tw, tid, L, n / 8, n / (L * 8));
			L *= 8;
		} else if (rem == 4) {
			fft_pass_rt<4, NT, MAXN, INV>(buf, tw, tid, L, n / 4, 1);
			L *= 4;
		} else {
			fft_pass_rt<2, NT, MAXN, INV>(buf, tw, tid, L, n / 2, 1);
			L *= 2;
		}
	}
}

// Bluestein DFT of length n held in logical buf[0..n); buf holds m logical
// entries (m a power of two >= 2n-1).  chirp[k] = exp(-pi i (k^2 mod 2n) / n);
// bhat = FFT_m of b[j] = conj(chirp[|j|]) (wrapped).  twm: the m-point
// twiddle table.  The inverse uses conj(chirp) and
// bhat_inv[j] = conj(bhat[(m-j) mod m]).
template <int NT, int MAXN, bool INV>
__device__ __forceinline__ void bluestein_lds(float2* buf, int n, int m, const float2* __restrict__ chirp,
	const float2* __restrict__ bhat, const float2* __restrict__ twm, int tid)
{
	for (int k = tid; k < m; k += NT) {
		if (k < n) {
			float2 w = chirp[k];
			if (INV) w.y = -w.y;
			buf[pidx(k)] = cmul(buf[pidx(k)], w);
		} else {
			buf[pidx(k)] = make_float2(0.f, 0.f);
		}
	}
	__syncthreads();
	fft_lds_rt<NT, MAXN, false>(buf, m, twm, tid);
	for (int k = tid; k < m; k += NT) {
		float2 b;
		if (INV) {
			b = bhat[(m - k) & (m - 1)];
			b.y = -b.y;
		} else {
			b = bhat[k];
		}
		buf[pidx(k)] = cmul(buf[pidx(k)], b);
	}
	__syncthreads();
	fft_lds_rt<NT, MAXN, true>(buf, m, twm, tid);
	const float inv_m = 1.0f / (float)m;
	for (int k = tid; k < n; k += NT) {
		float2 w = chirp[k];
		if (INV) w.y = -w.y;
		float2 v = cmul(buf[pidx(k)], w);
		buf[pidx(k)] = make_float2(v.x * inv_m, v.y * inv_m);
	}
	__syncthreads();
}

// A 1D transform of arbitrary length n in LDS, as planned on the host.
struct Fft1dPlan {
	int n;                   // transform length
	int m;                   // working length: n for powers of two and for lengths 2^a 3^b 5^c,
	                         // else the Bluestein length, a power of two >= 2n-1
	const float2* tw;        // m-point twiddles
	const float2* chirp;     // n entries (Bluestein only, else null)
	const float2* bhat;      // m entries (Bluestein only, else null)
	int nrad;                // mixed-radix passes (lengths 2^a 3^b 5^c that are not powers of two), else 0
	unsigned char rad[20];   // their radices, in pass order: 8s, then a 4 or 2, then 3s, then 5s
	// Bluestein lengths n = 4 q whose quarter q has Bluestein length m / 4:
	// the transform may run as one radix-4 Cooley-Tukey step over four
	// q-point Bluestein transforms (ct4_bluestein); else sub_chirp is null
	const float2* sub_chirp; // q entries: the q-point chirp
	const float2* sub_bhat;  // m / 4 entries
	const float2* sub_tw;    // m / 4-point twiddles
	const float2* ct_tw;     // n entries: exp(-2 pi i e / n)
};

// ---- mixed-radix FFT (lengths 2^a 3^b 5^c) --------------------------------
// The Stockham pass of radix R after passes of product L, for any radices:
// butterfly j (T = n / R of them) reads j + r T, multiplies element r by
// w^r with w = exp(-+2 pi i k S / n), k = j mod L, S = n / (L R), and
// writes (j - k) R + k + r L.
template <int R, int NT, int MAXN, bool INV>
__device__ __forceinline__ void fft_pass_mx(float2* buf, const float2* __restrict__ tw, int tid,
	int L, int T, int S)
{
	constexpr int BPT = (MAXN / R + NT - 1) / NT;
	const bool lpow2 = (L & (L - 1)) == 0;
	float2 u[BPT][R];
	int kk[BPT];
#pragma unroll
	for (int b = 0; b < BPT; ++b) {
		const int j = tid + b * NT;
		kk[b] = 0;
		if (j < T) {
			kk[b] = lpow2 ? (j & (L - 1)) : j % L;
#pragma unroll
			for (int r = 0; r < R; ++r) u[b][r] = buf[pidx(j + r * T)];
			if (L > 1) {
				float2 w = tw[kk[b] * S];
				if (INV) w.y = -w.y;
				twiddle_powers<R>(u[b], w);
			}
			dft_small<R, INV>(u[b]);
		}
	}
	__syncthreads();
#pragma unroll
	for (int b = 0; b < BPT; ++b) {
		const int j = tid + b * NT;
		if (j < T) {
			const int base = (j - kk[b]) * R + kk[b];
#pragma unroll
			for (int r = 0; r < R; ++r) buf[pidx(base + r * L)] = u[b][r];
		}
	}
	__syncthreads();
}

// In-place transform of a mixed-radix plan (n <= MAXN); same contract as
// fft_lds_rt.
template <int NT, int MAXN, bool INV>
__device__ __forceinline__ void fft_lds_mixed(float2* buf, const Fft1dPlan& p, int tid)
{
	int L = 1;
	for (int i = 0; i < p.nrad; ++i) {
		const int R = p.rad[i];
		const int T = p.n / R, S = p.n / (L * R);
		if (R == 8) fft_pass_mx<8, NT, MAXN, INV>(buf, p.tw, tid, L, T, S);
		else if (R == 4) fft_pass_mx<4, NT, MAXN, INV>(buf, p.tw, tid, L, T, S);
		else if (R == 2) fft_pass_mx<2, NT, MAXN, INV>(buf, p.tw, tid, L, T, S);
		else if (R == 3) fft_pass_mx<3, NT, MAXN, INV>(buf, p.tw, tid, L, T, S);
		else fft_pass_mx<5, NT, MAXN, INV>(buf, p.tw, tid, L, T, S);
		L *= R;
	}
}

// ---- one radix-4 step over Bluestein transforms ---------------------------
// For n = 4 q: X[k1 + q k2] = sum_g W^(g k1) Y_g[k1] exp(-+2 pi i g k2 / 4),
// W = exp(-+2 pi i / n), with Y_g the q-point DFT of x[4 m + g].  Thread
// group g (NT/4 threads) runs Y_g as a Bluestein transform on M/4 points in
// its quarter of buf; the four read their inputs (src, which may read buf)
// in the first pass, before its barrier, so no quarter is written before
// every input is read.  The combine reads all Y into registers before the
// barrier that precedes its writes.  Needs NT/4 > 64 (barriers span the
// workgroup) and MAXN/4 = the plan's m/4.  Result in logical buf[0..n).
template <int NT, int MAXN, bool INV, class Src>
__device__ __forceinline__ void ct4_bluestein(float2* buf, const Fft1dPlan& p, int tid, const Src& src)
{
	constexpr int SNT = NT / 4, SM = MAXN / 4;
	static_assert(SNT > 64, "the sub-transforms' barriers must span the workgroup");
	const int g = tid / SNT, lt = tid - g * SNT;
	const int q = p.n / 4;
	float2* sb = buf + g * fft_lds_len(SM);           // = buf + pidx(g SM): SM is a multiple of 8
	auto ld = [&](int i) -> float2 { return src(4 * i + g); };
	bluestein_io<SNT, SM, INV>(sb, q, p.sub_chirp, p.sub_bhat, p.sub_tw, lt, ld);
	constexpr int PER = (MAXN / 4 + NT - 1) / NT;
	float2 y[PER][4];
#pragma unroll
	for (int b = 0; b < PER; ++b) {
		const int t = tid + b * NT;
		if (t < q) {
#pragma unroll
			for (int h = 0; h < 4; ++h) {
				const float2 v = buf[h * fft_lds_len(SM) + pidx(t)];
				if (h == 0) {
					y[b][h] = v;
				} else {
					const float2 w = p.ct_tw[h * t];
					y[b][h] = INV ? cmulc(v, w) : cmul(v, w);
				}
			}
			dft4<INV>(y[b][0], y[b][1], y[b][2], y[b][3]);
		}
	}
	__syncthreads();
#pragma unroll
	for (int b = 0; b < PER; ++b) {
		const int t = tid + b * NT;
		if (t < q) {
#pragma unroll
			for (int h = 0; h < 4; ++h) buf[pidx(t + h * q)] = y[b][h];
		}
	}
	__syncthreads();
}

template <int NT, int MAXN, bool INV>
__device__ __forceinline__ void fft_any_lds(float2* buf, const Fft1dPlan& p, int tid)
{
	if (p.nrad > 0) {
		fft_lds_mixed<NT, MAXN, INV>(buf, p, tid);
	} else if (p.chirp == nullptr) {
		fft_lds_rt<NT, MAXN, INV>(buf, p.n, p.tw, tid);
	} else if (p.m == MAXN) {
		if constexpr (NT / 4 > 64) {
			if (p.sub_chirp != nullptr) {
				ct4_bluestein<NT, MAXN, INV>(buf, p, tid, LdsIn{buf});
				return;
			}
		}
		bluestein_io<NT, MAXN, INV>(buf, p.n, p.chirp, p.bhat, p.tw, tid, LdsIn{buf});
	} else {
		bluestein_lds<NT, MAXN, INV>(buf, p.n, p.m, p.chirp, p.bhat, p.tw, tid);
	}
}

// The same transform with the input element i < p.n given by src(i) instead
// of buf (src may read buf).  Bluestein plans of working length MAXN read src
// inside the first pass; other plans first gather src into buf.
template <int NT, int MAXN, bool INV, class Src>
__device__ __forceinline__ void fft_any_lds_src(float2* buf, const Fft1dPlan& p, int tid, const Src& src)
{
	if (p.chirp != nullptr && p.m == MAXN) {
		if constexpr (NT / 4 > 64) {
			if (p.sub_chirp != nullptr) {
				ct4_bluestein<NT, MAXN, INV>(buf, p, tid, src);
				return;
			}
		}
		bluestein_io<NT, MAXN, INV>(buf, p.n, p.chirp, p.bhat, p.tw, tid, src);
		return;
	}
	constexpr int PER = (MAXN + NT - 1) / NT;
	float2 v[PER];
#pragma unroll
	for (int i = 0; i < PER; ++i) {
		const int y = tid + i * NT;
		if (y < p.n) v[i] = src(y);
	}
	__syncthreads();
#pragma unroll
	for (int i = 0; i < PER; ++i) {
		const int y = tid + i * NT;
		if (y < p.n) buf[pidx(y)] = v[i];
	}
	__syncthreads();
	fft_any_lds<NT, MAXN, INV>(buf, p, tid);
}

}  // namespace gctf
