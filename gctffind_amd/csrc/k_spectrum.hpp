// k_spectrum.hpp -- spectrum-stage kernels: Fourier-crop rescale, tiled
// amplitude spectrum, 2D background removal, flatness gate, low-pass and
// radial average.  Each kernel cites the reference function it replaces.
#pragma once
#include "gctf_internal.hpp"

namespace gctf {

__device__ __forceinline__ float load_px(const float* p, long long i) { return p[i]; }
__device__ __forceinline__ float load_px(const _Float16* p, long long i) { return (float)p[i]; }

// ---------------------------------------------------------------------------
// Rescale (CRescaleImage.cpp:47-99; CCufft2D.cpp:23-114; GFtResize2D.cu:7-32)
// ---------------------------------------------------------------------------

// The 8 XCDs of the chip take workgroups round-robin (XCD = linear id mod
// 8).  xcd_group(b, n) maps a grid of 8 ceil(n / 8) workgroups so that each
// XCD works through one contiguous range of [0, n): neighbouring rows or
// columns then meet in the same L2, which merges their partial-line writes.
// Returns -1 for the padding workgroups.
__device__ __forceinline__ int xcd_group(int b, int n)
{
	const int per = (n + 7) / 8;
	const int i = (b & 7) * per + (b >> 3);
	return i < n ? i : -1;
}

// Row stride (complex elements) of the row spectra the R2C pass writes for
// the transpose: ncx rounded up to 16, so that every row starts on a
// 128-byte line and the transpose's 32-element tile rows (256 bytes) read
// two whole lines each.
__host__ __device__ constexpr int spec_ld(int ncx) { return (ncx + 15) & ~15; }

struct TileGrid {
	int ntx, nty, offx, offy, step, tile;
	int blocks;          // 1: tile sums from 2x2 blocks of step x step
	int nbx, nby;        // block counts
	int bsize;           // block side (step, or tile when blocks == 0)
};

// One micrograph of a tile work list: the tile passes of an uneven list of
// micrographs run as single launches over all their work items, each
// workgroup finding its micrograph here.  Offsets are in elements of the
// shared buffers.
struct TileJob {
	const void* img;          // pixels the tile passes read (the launch's Tin)
	int pitch;                // row pitch in pixels
	TileGrid g;
	int rows_first;           // first row-pass item (prefix over the launch's jobs)
	int nsplit;               // tile-column splits
	int need_rows;            // 1: row partials still to be summed from img
	int raw_slot;             // spectrum slot of the micrograph
	long long rowsum_off, sums_off, rt_off, part_off;
};

// The job holding row-pass item `item`: the last with rows_first <= item.
__device__ __forceinline__ int find_job(const TileJob* __restrict__ jobs, int njobs, int item)
{
	int lo = 0, hi = njobs - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (jobs[mid].rows_first <= item) lo = mid;
		else hi = mid - 1;
	}
	return lo;
}

// Sum over a wavefront of the lanes' partial sums of a row segment (xor
// butterfly, a fixed order); every lane returns the same value.
__device__ __forceinline__ float wave_row_sum(float v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
	return v;
}

// Waves per SIMD the rescale kernels are compiled for: their LDS buffer
// (fft_lds_len(MAXN) complex values) leaves room for 32 waves per CU up to
// MAXN = 8192.
__host__ __device__ constexpr int rescale_waves(int maxn) { return maxn <= 8192 ? 8 : 4; }

// Real rows -> half-spectrum rows kept for kx < ncx_out, spec_ld(ncx_out)
// apart.  Two rows share one complex transform (z = a + i b).  Row pair bx
// of one image.
template <typename Tin, int NT, int MAXN>
__device__ __forceinline__ void rows_r2c_body(const Tin* __restrict__ img, int nx, int ny, const Fft1dPlan& plan,
	int ncx_out, float2* __restrict__ out, int bx)
{
	__shared__ float2 buf[fft_lds_len(MAXN)];
	const int tid = threadIdx.x;
	const int y0 = 2 * bx;
	const bool two = (y0 + 1) < ny;
	const Tin* src = img + (long long)y0 * nx;
	for (int x = tid; x < nx; x += NT) {
		float a = load_px(src, x);
		float b = two ? load_px(src, (long long)nx + x) : 0.f;
		buf[pidx(x)] = make_float2(a, b);
	}
	__syncthreads();
	fft_any_lds<NT, MAXN, false>(buf, plan, tid);
	const int ld = spec_ld(ncx_out);
	float2* dst = out + (long long)y0 * ld;
	for (int k = tid; k < ncx_out; k += NT) {
		float2 z = buf[pidx(k)];
		float2 zc = buf[pidx(k == 0 ? 0 : nx - k)];
		// A = (Z[k] + conj Z[N-k]) / 2 ; B = -i (Z[k] - conj Z[N-k]) / 2
		float2 A = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y - zc.y));
		float2 B = make_float2(0.5f * (z.y + zc.y), -0.5f * (z.x - zc.x));
		dst[k] = A;
		if (two) dst[ld + k] = B;
	}
}

// Grid (ceil(ny/2), images).
template <typename Tin, int NT, int MAXN>
__global__ __launch_bounds__(NT, rescale_waves(MAXN)) void k_rescale_rows_r2c(const Tin* __restrict__ img,
	long long img_stride, int nx, int ny, Fft1dPlan plan, int ncx_out,
	float2* __restrict__ out, long long out_stride)
{
	rows_r2c_body<Tin, NT, MAXN>(img + blockIdx.y * img_stride, nx, ny, plan, ncx_out,
		out + blockIdx.y * out_stride, blockIdx.x);
}

// Rows too long for one LDS transform of length nx (nx even, not a power of
// two, Bluestein length above MAXN): one row per workgroup as a complex
// transform of length h = nx/2 of z[m] = x[2m] + i x[2m+1].  With
// E = (Z[k] + conj Z[h-k]) / 2 and O = -i (Z[k] - conj Z[h-k]) / 2 the
// transforms of the even and odd samples, X[k] = E[k] + exp(-2 pi i k/nx) O[k]
// for k < ncx_out <= h, rows spec_ld(ncx_out) apart.  plan is the length-h
// plan.  Row y of one image.
template <typename Tin, int NT, int MAXN>
__device__ __forceinline__ void rows_r2c_half_body(const Tin* __restrict__ img, int nx, const Fft1dPlan& plan,
	int ncx_out, float2* __restrict__ out, int y)
{
	__shared__ float2 buf[fft_lds_len(MAXN)];
	const int tid = threadIdx.x;
	const int h = nx / 2;
	const Tin* src = img + (long long)y * nx;
	for (int m = tid; m < h; m += NT) buf[pidx(m)] = make_float2(load_px(src, 2 * m), load_px(src, 2 * m + 1));
	__syncthreads();
	fft_any_lds<NT, MAXN, false>(buf, plan, tid);
	float2* dst = out + (long long)y * spec_ld(ncx_out);
	for (int k = tid; k < ncx_out; k += NT) {
		const float2 z = buf[pidx(k)];
		const float2 zc = buf[pidx(k == 0 ? 0 : h - k)];
		const float2 E = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y - zc.y));
		const float2 O = make_float2(0.5f * (z.y + zc.y), -0.5f * (z.x - zc.x));
		float sn, cs;
		sincospif((float)k / (float)h, &sn, &cs);      // angle pi k / h = 2 pi k / nx
		dst[k] = make_float2(E.x + (cs * O.x + sn * O.y), E.y + (cs * O.y - sn * O.x));
	}
}

// Grid (ny, images).
template <typename Tin, int NT, int MAXN>
__global__ __launch_bounds__(NT, rescale_waves(MAXN)) void k_rescale_rows_r2c_half(const Tin* __restrict__ img,
	long long img_stride, int nx, int ny, Fft1dPlan plan, int ncx_out,
	float2* __restrict__ out, long long out_stride)
{
	(void)ny;
	rows_r2c_half_body<Tin, NT, MAXN>(img + blockIdx.y * img_stride, nx, plan, ncx_out,
		out + blockIdx.y * out_stride, blockIdx.x);
}

// Complex transpose out[x][y] = in[y][x]; in is rows x cols, rows
// spec_ld(cols) apart (the R2C pass's output).  32x32 tiles; tile (bx, by)
// of one image.
__device__ __forceinline__ void transpose_body(const float2* __restrict__ src, int rows, int cols,
	float2* __restrict__ dst, int bxi, int byi)
{
	const int ld = spec_ld(cols);
	__shared__ float2 t[32][33];
	const int bx = bxi * 32, by = byi * 32;
	const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // ty in [0, 8)
	for (int j = ty; j < 32; j += 8) {
		int y = by + j, x = bx + tx;
		if (y < rows && x < cols) t[j][tx] = src[(long long)y * ld + x];
	}
	__syncthreads();
	for (int j = ty; j < 32; j += 8) {
		int x = bx + j, y = by + tx;
		if (x < cols && y < rows) dst[(long long)x * rows + y] = t[tx][j];
	}
}

__global__ __launch_bounds__(256) void k_transpose_c(const float2* __restrict__ in,
	long long in_stride, int rows, int cols, float2* __restrict__ out, long long out_stride)
{
	transpose_body(in + blockIdx.z * in_stride, rows, cols, out + blockIdx.z * out_stride, blockIdx.x, blockIdx.y);
}

// Columns (rows of the transposed spectrum in[kx][y]): forward length ny,
// crop to nny rows (ky <= nny/2 from the top, the rest from the bottom),
// inverse length nny, written back row-major: out[y'][kx] (ncx entries per
// y').  Column kx of one image.
template <int NT, int MAXN>
__device__ __forceinline__ void cols_body(const float2* __restrict__ in, int ny, int nny, int ncx,
	const Fft1dPlan& pf, const Fft1dPlan& pi, float2* __restrict__ out, int kx)
{
	__shared__ float2 buf[fft_lds_len(MAXN)];
	const int tid = threadIdx.x;
	const float2* src = in + (long long)kx * ny;
	for (int y = tid; y < ny; y += NT) buf[pidx(y)] = src[y];
	__syncthreads();
	fft_any_lds<NT, MAXN, false>(buf, pf, tid);
	// the crop is the inverse transform's input
	auto crop = [&](int y) -> float2 {
		const int ys = y > nny / 2 ? y - nny + ny : y;
		return buf[pidx(ys)];
	};
	fft_any_lds_src<NT, MAXN, true>(buf, pi, tid, crop);
	float2* dst = out + kx;
	for (int y = tid; y < nny; y += NT) dst[(long long)y * ncx] = buf[pidx(y)];
}

// Grid (8 ceil(ncx / 8), images).
template <int NT, int MAXN>
__global__ __launch_bounds__(NT, rescale_waves(MAXN)) void k_rescale_cols(const float2* __restrict__ in,
	long long in_stride, int ny, int nny, int ncx, Fft1dPlan pf, Fft1dPlan pi,
	float2* __restrict__ out, long long out_stride)
{
	const int kx = xcd_group(blockIdx.x, ncx);
	if (kx < 0) return;
	cols_body<NT, MAXN>(in + blockIdx.y * in_stride, ny, nny, ncx, pf, pi, out + blockIdx.y * out_stride, kx);
}

// Columns too long for one LDS transform (ny even, not a power of two, above
// MAXN/2; portrait super-resolution images): the forward transform of length
// ny is split into the transforms of the even and odd samples (length
// h = ny/2), X[k] = E[k mod h] + exp(-2 pi i k/ny) O[k mod h].  Each thread
// parks the E values of its kept rows y' in the output column (out[y'][kx])
// until O is done, and reads them back itself.  pf is the length-h plan;
// nny <= MAXN/2.  Column kx of one image.
template <int NT, int MAXN>
__device__ __forceinline__ void cols_split_body(const float2* __restrict__ in, int ny, int nny, int ncx,
	const Fft1dPlan& pf, const Fft1dPlan& pi, float2* __restrict__ out, int kx)
{
	constexpr int JN = (MAXN / 2 + NT - 1) / NT;     // kept rows per thread
	__shared__ float2 buf[fft_lds_len(MAXN)];
	const int tid = threadIdx.x;
	const int h = ny / 2;
	const float2* src = in + (long long)kx * ny;
	float2* dst = out + kx;
	for (int m = tid; m < h; m += NT) buf[pidx(m)] = src[2 * m];
	__syncthreads();
	fft_any_lds<NT, MAXN, false>(buf, pf, tid);
#pragma unroll
	for (int j = 0; j < JN; ++j) {
		const int y = tid + j * NT;
		if (y < nny) {
			int k = y > nny / 2 ? y - nny + ny : y;
			if (k >= h) k -= h;
			dst[(long long)y * ncx] = buf[pidx(k)];
		}
	}
	__syncthreads();
	for (int m = tid; m < h; m += NT) buf[pidx(m)] = src[2 * m + 1];
	__syncthreads();
	fft_any_lds<NT, MAXN, false>(buf, pf, tid);
	float2 x[JN];
#pragma unroll
	for (int j = 0; j < JN; ++j) {
		const int y = tid + j * NT;
		if (y < nny) {
			const int k = y > nny / 2 ? y - nny + ny : y;
			const float2 e = dst[(long long)y * ncx];
			const float2 o = buf[pidx(k >= h ? k - h : k)];
			float sn, cs;
			sincospif((float)k / (float)h, &sn, &cs);  // angle pi k / h = 2 pi k / ny
			x[j] = make_float2(e.x + (cs * o.x + sn * o.y), e.y + (cs * o.y - sn * o.x));
		}
	}
	__syncthreads();
#pragma unroll
	for (int j = 0; j < JN; ++j) {
		const int y = tid + j * NT;
		if (y < nny) buf[pidx(y)] = x[j];
	}
	__syncthreads();
	fft_any_lds<NT, MAXN, true>(buf, pi, tid);
	for (int y = tid; y < nny; y += NT) dst[(long long)y * ncx] = buf[pidx(y)];
}

// Grid (8 ceil(ncx / 8), images).
template <int NT, int MAXN>
__global__ __launch_bounds__(NT, rescale_waves(MAXN)) void k_rescale_cols_split(const float2* __restrict__ in,
	long long in_stride, int ny, int nny, int ncx, Fft1dPlan pf, Fft1dPlan pi,
	float2* __restrict__ out, long long out_stride)
{
	const int kx = xcd_group(blockIdx.x, ncx);
	if (kx < 0) return;
	cols_split_body<NT, MAXN>(in + blockIdx.y * in_stride, ny, nny, ncx, pf, pi, out + blockIdx.y * out_stride, kx);
}

// Half-spectrum rows -> real rows of length nnx (C2R; imaginary parts of the
// DC and Nyquist bins are dropped, i.e. the real part of the inverse of the
// Hermitian extension).  Two rows per transform.  Output rows have `pitch`
// floats; the pad columns are zeroed.  Row pair bx of one image.
//
// With rowsum, the row segments of the tile grid's mean blocks are summed
// too (one wavefront per block column, FP32, fixed order):
// rowsum[y - offy][bi] for the block rows y - offy in
// [0, (nby-1) step + bsize).  k_block_rows turns them into block sums.
template <int NT, int MAXN>
__device__ __forceinline__ void rows_c2r_body(const float2* __restrict__ in, int nny, int nnx,
	const Fft1dPlan& plan, float* __restrict__ out, int pitch, const TileGrid& g, float* __restrict__ rowsum, int bx)
{
	__shared__ float2 buf[fft_lds_len(MAXN)];
	const int tid = threadIdx.x;
	const int ncx = nnx / 2 + 1;
	const int y0 = 2 * bx;
	const bool two = (y0 + 1) < nny;
	const float2* ga = in + (long long)y0 * ncx;
	const float2* gb = ga + ncx;
	// Z = A + i B of the Hermitian extensions, read straight into the
	// transform's first pass
	auto herm = [&](int k) -> float2 {
		float2 a, b;
		if (k < ncx) {
			a = ga[k];
			b = two ? gb[k] : make_float2(0.f, 0.f);
			if (k == 0 || k == nnx / 2) { a.y = 0.f; b.y = 0.f; }
		} else {
			a = ga[nnx - k]; a.y = -a.y;
			b = two ? gb[nnx - k] : make_float2(0.f, 0.f);
			b.y = -b.y;
		}
		return make_float2(a.x - b.y, a.y + b.x);
	};
	fft_any_lds_src<NT, MAXN, true>(buf, plan, tid, herm);
	float* da = out + (long long)y0 * pitch;
	for (int x = tid; x < pitch; x += NT) {
		float2 z = x < nnx ? buf[pidx(x)] : make_float2(0.f, 0.f);
		da[x] = z.x;
		if (two) da[pitch + x] = z.y;
	}
	if (rowsum == nullptr) return;
	const int nr = (g.nby - 1) * g.step + g.bsize;
	const int ra = y0 - g.offy;
	if (ra + 1 < 0 || ra >= nr) return;
	const int wave = tid >> 6, lane = tid & 63;
	for (int bi = wave; bi < g.nbx; bi += NT / 64) {
		const int x0 = g.offx + bi * g.step;
		float sa = 0.f, sb = 0.f;
		for (int i = lane; i < g.bsize; i += 64) {
			const float2 z = buf[pidx(x0 + i)];
			sa += z.x;
			sb += z.y;
		}
		sa = wave_row_sum(sa);
		sb = wave_row_sum(sb);
		if (lane == 0) {
			if (ra >= 0) rowsum[ra * g.nbx + bi] = sa;
			if (two && ra + 1 < nr) rowsum[(ra + 1) * g.nbx + bi] = sb;
		}
	}
}

// Grid (ceil(nny/2), images); rowsum [image][block rows][nbx] or null.
template <int NT, int MAXN>
__global__ __launch_bounds__(NT, rescale_waves(MAXN)) void k_rescale_rows_c2r(const float2* __restrict__ in,
	long long in_stride, int nny, int nnx, Fft1dPlan plan, float* __restrict__ out,
	long long out_stride, int pitch, TileGrid g, float* __restrict__ rowsum)
{
	float* rs = rowsum == nullptr ? nullptr
		: rowsum + (long long)blockIdx.y * ((g.nby - 1) * g.step + g.bsize) * g.nbx;
	rows_c2r_body<NT, MAXN>(in + blockIdx.y * in_stride, nny, nnx, plan, out + blockIdx.y * out_stride, pitch, g,
		rs, blockIdx.x);
}

// C2R rows too long for one LDS transform of length nnx (as for the R2C rows
// above): one row per workgroup.  With h = nnx/2 and X the Hermitian
// extension of the half row, the inverse transform of length h of
// F[k] = (X[k] + X[k+h]) + i exp(2 pi i k/nnx) (X[k] - X[k+h]) is
// z[m] = x[2m] + i x[2m+1].  plan is the length-h plan.  Row partials as in
// k_rescale_rows_c2r (same pixels, same order).  Row y of one image.
template <int NT, int MAXN>
__device__ __forceinline__ void rows_c2r_half_body(const float2* __restrict__ in, int nnx, const Fft1dPlan& plan,
	float* __restrict__ out, int pitch, const TileGrid& g, float* __restrict__ rowsum, int y)
{
	__shared__ float2 buf[fft_lds_len(MAXN)];
	const int tid = threadIdx.x;
	const int ncx = nnx / 2 + 1, h = nnx / 2;
	const float2* ga = in + (long long)y * ncx;
	auto fold = [&](int k) -> float2 {
		float2 a = ga[k];                              // X[k], k < h
		float2 b = ga[h - k];                          // X[k+h] = conj X[h-k]
		b.y = -b.y;
		if (k == 0) { a.y = 0.f; b.y = 0.f; }          // X[0], X[h] real
		const float2 su = make_float2(a.x + b.x, a.y + b.y);
		const float2 di = make_float2(a.x - b.x, a.y - b.y);
		float sn, cs;
		sincospif((float)k / (float)h, &sn, &cs);      // exp(i pi k / h)
		const float2 t = make_float2(cs * di.x - sn * di.y, cs * di.y + sn * di.x);
		return make_float2(su.x - t.y, su.y + t.x);    // su + i t
	};
	fft_any_lds_src<NT, MAXN, true>(buf, plan, tid, fold);
	float* da = out + (long long)y * pitch;
	for (int x = tid; x < pitch; x += NT) {
		float v = 0.f;
		if (x < nnx) {
			const float2 z = buf[pidx(x >> 1)];
			v = (x & 1) ? z.y : z.x;
		}
		da[x] = v;
	}
	if (rowsum == nullptr) return;
	const int nr = (g.nby - 1) * g.step + g.bsize;
	const int ra = y - g.offy;
	if (ra < 0 || ra >= nr) return;
	const int wave = tid >> 6, lane = tid & 63;
	for (int bi = wave; bi < g.nbx; bi += NT / 64) {
		const int x0 = g.offx + bi * g.step;
		float sa = 0.f;
		for (int i = lane; i < g.bsize; i += 64) {
			const int x = x0 + i;
			const float2 z = buf[pidx(x >> 1)];
			sa += (x & 1) ? z.y : z.x;
		}
		sa = wave_row_sum(sa);
		if (lane == 0) rowsum[ra * g.nbx + bi] = sa;
	}
}

// Grid (nny, images).
template <int NT, int MAXN>
__global__ __launch_bounds__(NT, rescale_waves(MAXN)) void k_rescale_rows_c2r_half(const float2* __restrict__ in,
	long long in_stride, int nny, int nnx, Fft1dPlan plan, float* __restrict__ out,
	long long out_stride, int pitch, TileGrid g, float* __restrict__ rowsum)
{
	(void)nny;
	float* rs = rowsum == nullptr ? nullptr
		: rowsum + (long long)blockIdx.y * ((g.nby - 1) * g.step + g.bsize) * g.nbx;
	rows_c2r_half_body<NT, MAXN>(in + blockIdx.y * in_stride, nnx, plan, out + blockIdx.y * out_stride, pitch, g,
		rs, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Rescale over a work list of micrographs of different shapes, types and
// pixel sizes: each stage runs as one launch per kernel class (transform
// variant and length) over the workgroup items of every micrograph of that
// class, instead of one launch per micrograph.  The bodies are the
// single-image kernels', so both give the same bits.
// ---------------------------------------------------------------------------
struct RescaleJob {
	const void* img;          // raw pixels (the class launch's Tin)
	int nx, ny;               // raw size
	int nnx, nny, ncx;        // rescaled size, kept columns nnx/2 + 1
	int pitch;                // padded output row (floats)
	Fft1dPlan rows, cf, ci, c2r;
	float2* A;                // ny row spectra spec_ld(ncx) apart, then nny x ncx rescaled columns
	float2* B;                // ncx x ny transposed spectrum
	float* out;               // padded rescaled image
	float* rowsum;            // mean-block row partials of g (null: none)
	TileGrid g;
};

// Launch item `item` belongs to entry e with first[e] <= item < first[e+1];
// entry e is micrograph job[e].
__device__ __forceinline__ int find_entry(const int* __restrict__ first, int n, int item)
{
	int lo = 0, hi = n - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (first[mid] <= item) lo = mid;
		else hi = mid - 1;
	}
	return lo;
}

// Rows R2C (pairs, or half-length rows with HALF): items ceil(ny/2) or ny.
template <typename Tin, int NT, int MAXN, bool HALF>
__global__ __launch_bounds__(NT, rescale_waves(MAXN)) void k_rescale_rows_r2c_list(
	const RescaleJob* __restrict__ jobs, const int* __restrict__ first, const int* __restrict__ job, int n)
{
	const int e = find_entry(first, n, blockIdx.x);
	const RescaleJob& J = jobs[job[e]];
	const int b = blockIdx.x - first[e];
	if constexpr (HALF)
		rows_r2c_half_body<Tin, NT, MAXN>((const Tin*)J.img, J.nx, J.rows, J.ncx, J.A, b);
	else
		rows_r2c_body<Tin, NT, MAXN>((const Tin*)J.img, J.nx, J.ny, J.rows, J.ncx, J.A, b);
}

// Transpose A (ny x ncx) -> B: items ceil(ncx/32) ceil(ny/32).
__global__ __launch_bounds__(256) void k_transpose_list(const RescaleJob* __restrict__ jobs,
	const int* __restrict__ first, const int* __restrict__ job, int n)
{
	const int e = find_entry(first, n, blockIdx.x);
	const RescaleJob& J = jobs[job[e]];
	const int b = blockIdx.x - first[e], tx = (J.ncx + 31) / 32;
	transpose_body(J.A, J.ny, J.ncx, J.B, b % tx, b / tx);
}

// Columns B -> A (SPLIT: even/odd forward halves): items 8 ceil(ncx/8), so
// every micrograph's first item is a multiple of 8 and xcd_group sees the
// hardware's XCD of the workgroup.
template <int NT, int MAXN, bool SPLIT>
__global__ __launch_bounds__(NT, rescale_waves(MAXN)) void k_rescale_cols_list(
	const RescaleJob* __restrict__ jobs, const int* __restrict__ first, const int* __restrict__ job, int n)
{
	const int e = find_entry(first, n, blockIdx.x);
	const RescaleJob& J = jobs[job[e]];
	const int kx = xcd_group(blockIdx.x - first[e], J.ncx);
	if (kx < 0) return;
	if constexpr (SPLIT)
		cols_split_body<NT, MAXN>(J.B, J.ny, J.nny, J.ncx, J.cf, J.ci, J.A, kx);
	else
		cols_body<NT, MAXN>(J.B, J.ny, J.nny, J.ncx, J.cf, J.ci, J.A, kx);
}

// Rows C2R A -> padded image with the mean-block row partials (HALF:
// half-length rows): items ceil(nny/2) or nny.
template <int NT, int MAXN, bool HALF>
__global__ __launch_bounds__(NT, rescale_waves(MAXN)) void k_rescale_rows_c2r_list(
	const RescaleJob* __restrict__ jobs, const int* __restrict__ first, const int* __restrict__ job, int n)
{
	const int e = find_entry(first, n, blockIdx.x);
	const RescaleJob& J = jobs[job[e]];
	const int b = blockIdx.x - first[e];
	if constexpr (HALF)
		rows_c2r_half_body<NT, MAXN>(J.A, J.nnx, J.c2r, J.out, J.pitch, J.g, J.rowsum, b);
	else
		rows_c2r_body<NT, MAXN>(J.A, J.nny, J.nnx, J.c2r, J.out, J.pitch, J.g, J.rowsum, b);
}

// ---------------------------------------------------------------------------
// Tile spectrum (CGenAvgSpectrum.cpp:84-139, GCalcMoment2D.cu, GNormalize2D.cu,
// GRoundEdge.cu, GCalcSpectrum.cu:15-31,90-105, GAddImages.cu)
// ---------------------------------------------------------------------------


// Tile means (CGenAvgSpectrum.cpp:84-139) come from FP64 sums of
// bsize x bsize blocks at (offx + i*step, offy + j*step), each block split
// into kSumParts row groups: sums[image][block][part].  They are built in
// two steps.  First the row partials: for every block row y - offy in
// [0, (nby-1) step + bsize) and block column bi, the FP32 sum of the row's
// bsize pixels, summed by one wavefront (lane l adds pixels l, l+64, ...,
// then an xor butterfly) into rowsum[image][y - offy][bi].  The rescale's
// C2R pass makes them from its output rows in LDS; k_row_sums makes them
// from an image in memory, in the same order, so both give the same bits.
constexpr int kSumParts = 8;

// One block row r per workgroup of 256 threads, one block column per
// wavefront.  img and rowsum are the micrograph's own.
template <typename Tin>
__device__ __forceinline__ void row_sums_body(const Tin* __restrict__ img, int pitch, const TileGrid& g,
	float* __restrict__ rowsum, int r)
{
	const Tin* src = img + (long long)(g.offy + r) * pitch;
	float* rs = rowsum + (long long)r * g.nbx;
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	// wave w sums blocks w, w + 4, ...: the loads of PB of them are issued
	// before any is reduced (each block's sum keeps its lane order and
	// butterfly, so the bits do not depend on PB)
	constexpr int PB = 4;
	for (int b0 = wave; b0 < g.nbx; b0 += 4 * PB) {
		float a[PB];
#pragma unroll
		for (int q = 0; q < PB; ++q) {
			const int bi = b0 + 4 * q;
			a[q] = 0.f;
			if (bi < g.nbx) {
				const int x0 = g.offx + bi * g.step;
#pragma unroll 4
				for (int i = lane; i < g.bsize; i += 64) a[q] += load_px(src, x0 + i);
			}
		}
#pragma unroll
		for (int q = 0; q < PB; ++q) {
			const int bi = b0 + 4 * q;
			const float t = wave_row_sum(a[q]);
			if (bi < g.nbx && lane == 0) rs[bi] = t;
		}
	}
}

// Grid (nr, images).
template <typename Tin>
__global__ __launch_bounds__(256) void k_row_sums(const Tin* __restrict__ img, long long img_stride,
	int pitch, TileGrid g, float* __restrict__ rowsum)
{
	const int nr = (g.nby - 1) * g.step + g.bsize;
	row_sums_body<Tin>(img + blockIdx.y * img_stride, pitch, g, rowsum + (long long)blockIdx.y * nr * g.nbx,
		blockIdx.x);
}

// Grid (largest nr, jobs): the jobs with need_rows.
template <typename Tin>
__global__ __launch_bounds__(256) void k_row_sums_list(const TileJob* __restrict__ jobs, float* __restrict__ rowsum)
{
	const TileJob& J = jobs[blockIdx.y];
	const int nr = (J.g.nby - 1) * J.g.step + J.g.bsize;
	if (!J.need_rows || (int)blockIdx.x >= nr) return;
	row_sums_body<Tin>((const Tin*)J.img, J.pitch, J.g, rowsum + J.rowsum_off, blockIdx.x);
}

// Second step: part p of block (bi, bj) adds its rows' partials in FP32,
// rows r0 + j into accumulator j % 4, then (a0 + a1) + (a2 + a3): a fixed
// order whose rounding error grows with a quarter of the row count.  Grid
// (ceil(nbx nby kSumParts / 256), images).
__device__ __forceinline__ void block_rows_body(const float* __restrict__ rs, const TileGrid& g,
	float* __restrict__ sums, int i)
{
	const int nblk = g.nbx * g.nby;
	if (i >= nblk * kSumParts) return;
	const int b = i / kSumParts, p = i % kSumParts;
	const int bi = b % g.nbx, bj = b / g.nbx;
	const int rows = (g.bsize + kSumParts - 1) / kSumParts;
	const int r0 = bj * g.step + p * rows, r1 = bj * g.step + min(g.bsize, (p + 1) * rows);
	float a[4] = {0.f, 0.f, 0.f, 0.f};
	for (int r = r0; r < r1; ++r) a[(r - r0) & 3] += rs[r * g.nbx + bi];
	sums[(long long)b * kSumParts + p] = (a[0] + a[1]) + (a[2] + a[3]);
}

__global__ __launch_bounds__(256) void k_block_rows(const float* __restrict__ rowsum, TileGrid g,
	float* __restrict__ sums)
{
	const int nr = (g.nby - 1) * g.step + g.bsize;
	block_rows_body(rowsum + (long long)blockIdx.y * nr * g.nbx, g,
		sums + (long long)blockIdx.y * g.nbx * g.nby * kSumParts, blockIdx.x * 256 + threadIdx.x);
}

// Grid (ceil(largest nbx nby kSumParts / 256), jobs).
__global__ __launch_bounds__(256) void k_block_rows_list(const TileJob* __restrict__ jobs,
	const float* __restrict__ rowsum, float* __restrict__ sums)
{
	const TileJob& J = jobs[blockIdx.y];
	block_rows_body(rowsum + J.rowsum_off, J.g, sums + J.sums_off, blockIdx.x * 256 + threadIdx.x);
}

// A block's sum from its kSumParts parts, pairwise in FP32.
__device__ __forceinline__ float block_sum(const float* bs, int b)
{
	const float* q = bs + b * kSumParts;
	static_assert(kSumParts == 8, "pairwise order written for 8 parts");
	return ((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + q[7]));
}

// Row pass of the tile FFT: for ROWS rows of one tile, subtract the tile
// mean, apply the round-edge mask, R2C transform along x (two rows per
// complex transform) and write the half-spectrum transposed:
// rt[image][tile][kx][y].  A workgroup of (ROWS/2)*(T/8) threads takes G
// consecutive row blocks (block b = rows [(b % (T/ROWS)) ROWS, +ROWS) of
// tile b / (T/ROWS)); the image rows of the next block are loaded into
// registers while the current one is transformed.  Grid
// (ceil(ntiles * T/ROWS / G), images).  The transposing stage reuses the
// transforms' LDS (the split halves wait in registers across a barrier).
// The pass for item `item` (row blocks item*G ..) of one micrograph: im, bs
// and rt are the micrograph's pixels, block sums and row-pass output.
template <typename Tin, int T, int ROWS, int G>
__device__ __forceinline__ void tile_rows_body(const Tin* __restrict__ im, int pitch, const TileGrid& g,
	const float* __restrict__ bs, const float* __restrict__ mask, const float2* __restrict__ tw,
	float2* __restrict__ rt, int item)
{
	constexpr int FT = T / 8;              // threads per transform
	constexpr int NG = ROWS / 2;           // transforms per workgroup
	constexpr int NT = NG * FT;
	constexpr int CX = T / 2 + 1;
	constexpr int FL = fft_lds_len(T);
	constexpr int PX = T / FT;             // pixels per thread and row
	constexpr int KP = (CX + FT - 1) / FT; // half-spectrum entries per thread
	constexpr int SP = ROWS + 1;           // stage pitch: lanes of a wave write
	                                       // k-consecutive entries, an odd pitch
	                                       // keeps them off one LDS bank
	constexpr int NRB = T / ROWS;          // row blocks per tile
	static_assert(CX * SP <= NG * FL, "stage must fit in the transform buffers");
	__shared__ float2 buf[NG * FL];
	float2* stage = buf;
	const int nblk = g.ntx * g.nty * NRB;
	const int grp = threadIdx.x / FT, lane = threadIdx.x % FT;
	float2* fb = buf + grp * FL;
	// this thread's pixels of the two rows of row block b, and their mask
	float va[PX], vb[PX], ma[PX], mb[PX];
	auto load_rows = [&](int b) {
		const int t = b / NRB, ya = (b % NRB) * ROWS + 2 * grp;
		const int sx = g.offx + (t % g.ntx) * g.step, sy = g.offy + (t / g.ntx) * g.step;
		const Tin* src = im + (long long)(sy + ya) * pitch + sx;
#pragma unroll
		for (int i = 0; i < PX; ++i) {
			va[i] = load_px(src, lane + i * FT);
			vb[i] = load_px(src, (long long)pitch + lane + i * FT);
		}
	};
	auto load_mask = [&](int b) {
		const int ya = (b % NRB) * ROWS + 2 * grp;
#pragma unroll
		for (int i = 0; i < PX; ++i) {
			ma[i] = mask[ya * T + lane + i * FT];
			mb[i] = mask[(ya + 1) * T + lane + i * FT];
		}
	};
	const int b0 = item * G;
	if (b0 >= nblk) return;
	load_rows(b0);
	load_mask(b0);
	// Loads complete in issue order, and a wait after a branch that may
	// have issued loads counts them as issued: the next block's loads are
	// therefore unconditional (clamped to the last block) and issued after
	// the loads the transform waits for, so they stay in flight through it.
#pragma unroll
	for (int it = 0; it < G; ++it) {
		const int b = b0 + it;
		if (b >= nblk) break;
		const int nb = min(b + 1, nblk - 1);
		const int rb = b % NRB, t = b / NRB;
		const int tx = t % g.ntx, ty = t / g.ntx;
		float tsum;
		if (g.blocks) {
			tsum = (block_sum(bs, ty * g.nbx + tx) + block_sum(bs, ty * g.nbx + tx + 1))
				+ (block_sum(bs, (ty + 1) * g.nbx + tx) + block_sum(bs, (ty + 1) * g.nbx + tx + 1));
		} else {
			tsum = block_sum(bs, ty * g.nbx + tx);
		}
		const float mean = tsum / ((float)T * (float)T);
#pragma unroll
		for (int i = 0; i < PX; ++i) {
			const int x = lane + i * FT;
			fb[pidx(x)] = make_float2((va[i] - mean) * ma[i], (vb[i] - mean) * mb[i]);
		}
		fft_sync<FT>();
		TwRegs<T, FT> twr;
		fft_tw_load(twr, tw, lane);
		if (it + 1 < G) load_rows(nb);
		fft_lds<T, FT, false>(fb, twr, lane);
		float2 sa[KP], sb[KP];
#pragma unroll
		for (int i = 0; i < KP; ++i) {
			const int k = lane + i * FT;
			if (k < CX) {
				float2 z = fb[pidx(k)];
				float2 zc = fb[pidx((T - k) & (T - 1))];
				sa[i] = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y - zc.y));
				sb[i] = make_float2(0.5f * (z.y + zc.y), -0.5f * (z.x - zc.x));
			}
		}
		__syncthreads();
#pragma unroll
		for (int i = 0; i < KP; ++i) {
			const int k = lane + i * FT;
			if (k < CX) {
				stage[k * SP + 2 * grp] = sa[i];
				stage[k * SP + 2 * grp + 1] = sb[i];
			}
		}
		__syncthreads();
		float2* dst = rt + (long long)t * CX * T + rb * ROWS;
		for (int i = threadIdx.x; i < CX * ROWS; i += NT) {
			int k = i / ROWS, yy = i % ROWS;
			dst[(long long)k * T + yy] = stage[k * SP + yy];
		}
		if (it + 1 < G) load_mask(nb);
		// the stage is the transforms' buffer: read out before it is refilled
		__syncthreads();
	}
}

template <typename Tin, int T, int ROWS, int G>
__global__ __launch_bounds__((ROWS / 2) * (T / 8), G > 1 ? 6 : 8) void k_tile_rows(const Tin* __restrict__ img,
	long long img_stride, int pitch, TileGrid g, const float* __restrict__ sums,
	const float* __restrict__ mask, const float2* __restrict__ tw,
	float2* __restrict__ rt, long long rt_stride)
{
	tile_rows_body<Tin, T, ROWS, G>(img + blockIdx.y * img_stride, pitch, g,
		sums + blockIdx.y * (long long)g.nbx * g.nby * kSumParts, mask, tw, rt + blockIdx.y * rt_stride, blockIdx.x);
}

// The same over a tile work list: grid (total items of the jobs).
template <typename Tin, int T, int ROWS, int G>
__global__ __launch_bounds__((ROWS / 2) * (T / 8), G > 1 ? 6 : 8) void k_tile_rows_list(
	const TileJob* __restrict__ jobs, int njobs, const float* __restrict__ sums,
	const float* __restrict__ mask, const float2* __restrict__ tw, float2* __restrict__ rt)
{
	const int j = find_job(jobs, njobs, blockIdx.x);
	const TileJob& J = jobs[j];
	tile_rows_body<Tin, T, ROWS, G>((const Tin*)J.img, J.pitch, J.g, sums + J.sums_off, mask, tw, rt + J.rt_off,
		blockIdx.x - J.rows_first);
}

// Persistent form over a tile work list: `gridDim.x` workgroups take items
// 0 .. gridDim.x - 1 first, then pull the next item of the list from a work
// queue (*next, an atomic counter zeroed before the launch) until all
// `items` are done.  Same items and arithmetic as k_tile_rows_list.
template <typename Tin, int T, int ROWS, int G>
__global__ __launch_bounds__((ROWS / 2) * (T / 8), G > 1 ? 6 : 8) void k_tile_rows_queue(
	const TileJob* __restrict__ jobs, int njobs, int items, int* __restrict__ next,
	const float* __restrict__ sums, const float* __restrict__ mask, const float2* __restrict__ tw,
	float2* __restrict__ rt)
{
	__shared__ int s_item;
	int item = blockIdx.x;
	while (item < items) {
		const TileJob& J = jobs[find_job(jobs, njobs, item)];
		tile_rows_body<Tin, T, ROWS, G>((const Tin*)J.img, J.pitch, J.g, sums + J.sums_off, mask, tw, rt + J.rt_off,
			item - J.rows_first);
		__syncthreads();
		if (threadIdx.x == 0) s_item = (int)gridDim.x + atomicAdd(next, 1);
		__syncthreads();
		item = s_item;
	}
}

// Column pass: for CPW kx-columns and the tiles [t0, t1) of split s,
// transform along y, take |F| and accumulate tile * (1/ntiles).  Writes the
// DC-shifted partial average part[image][s][y][kx].  Grid
// (ceil(CX/CPW), nsplit, images), CPW*(T/8) threads.
// One micrograph: rt its row-pass output, part its split partials; column
// group cg, split s.
template <int T, int CPW>
__device__ __forceinline__ void tile_cols_body(const float2* __restrict__ src, int ntiles, int nsplit,
	float fac, const float2* __restrict__ tw, float* __restrict__ part, int cg, int s)
{
	constexpr int FT = T / 8;
	constexpr int CX = T / 2 + 1;
	constexpr int PER = T / FT;            // 8 outputs per thread
	constexpr int FL = fft_lds_len(T);
	__shared__ float2 buf[CPW * FL];
	const int grp = threadIdx.x / FT, lane = threadIdx.x % FT;
	const int kx = cg * CPW + grp;
	const bool live = kx < CX;
	const int per = (ntiles + nsplit - 1) / nsplit;
	const int t0 = s * per, t1 = min(ntiles, t0 + per);
	float acc[PER];
#pragma unroll
	for (int i = 0; i < PER; ++i) acc[i] = 0.f;
	float2* fb = buf + grp * FL;
	TwRegs<T, FT> twr;
	fft_tw_load(twr, tw, lane);
	for (int t = t0; t < t1; ++t) {
		if (live) {
			const float2* col = src + ((long long)t * CX + kx) * T;
			for (int y = lane; y < T; y += FT) fb[pidx(y)] = col[y];
		}
		fft_sync<FT>();
		fft_lds<T, FT, false>(fb, twr, lane);
#pragma unroll
		for (int i = 0; i < PER; ++i) {
			float2 c = fb[pidx(lane + i * FT)];
			acc[i] = fmaf(sqrtf(fmaf(c.x, c.x, c.y * c.y)), fac, acc[i]);
		}
		fft_sync<FT>();
	}
	if (!live) return;
	float* dst = part + (long long)s * T * CX;
#pragma unroll
	for (int i = 0; i < PER; ++i) {
		int y = lane + i * FT;
		int yd = (y + T / 2) & (T - 1);
		dst[(long long)yd * CX + kx] = acc[i];
	}
}

template <int T, int CPW>
__global__ __launch_bounds__(CPW * (T / 8)) void k_tile_cols(const float2* __restrict__ rt,
	long long rt_stride, int ntiles, int nsplit, float fac, const float2* __restrict__ tw,
	float* __restrict__ part)
{
	constexpr int CX = T / 2 + 1;
	tile_cols_body<T, CPW>(rt + blockIdx.z * rt_stride, ntiles, nsplit, fac, tw,
		part + (long long)blockIdx.z * nsplit * T * CX, blockIdx.x, blockIdx.y);
}

// The same over a tile work list: grid (ceil(CX/CPW), largest nsplit, jobs).
template <int T, int CPW>
__global__ __launch_bounds__(CPW * (T / 8)) void k_tile_cols_list(const TileJob* __restrict__ jobs,
	const float2* __restrict__ rt, const float2* __restrict__ tw, float* __restrict__ part)
{
	const TileJob& J = jobs[blockIdx.z];
	if ((int)blockIdx.y >= J.nsplit) return;
	const int ntiles = J.g.ntx * J.g.nty;
	tile_cols_body<T, CPW>(rt + J.rt_off, ntiles, J.nsplit, 1.0f / (float)ntiles, tw, part + J.part_off,
		blockIdx.x, blockIdx.y);
}

// Zero the work-queue counters of a list's persistent launches.
__global__ void k_queue_reset(int* __restrict__ q, int n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) q[i] = 0;
}

// Sum the split partials in split order: raw[image][i] = sum_s part[s][i].
__device__ __forceinline__ void tile_reduce_body(const float* __restrict__ p, int nsplit, int n,
	float* __restrict__ raw, int i)
{
	if (i >= n) return;
	float acc = 0.f;
	for (int s = 0; s < nsplit; ++s) acc += p[(long long)s * n + i];
	raw[i] = acc;
}

__global__ __launch_bounds__(256) void k_tile_reduce(const float* __restrict__ part, int nsplit,
	int n, float* __restrict__ raw)
{
	tile_reduce_body(part + (long long)blockIdx.y * nsplit * n, nsplit, n, raw + (long long)blockIdx.y * n,
		blockIdx.x * 256 + threadIdx.x);
}

// Grid (ceil(n / 256), jobs): raw slot per job.
__global__ __launch_bounds__(256) void k_tile_reduce_list(const TileJob* __restrict__ jobs,
	const float* __restrict__ part, int n, float* __restrict__ raw)
{
	const TileJob& J = jobs[blockIdx.y];
	tile_reduce_body(part + J.part_off, J.nsplit, n, raw + (long long)J.raw_slot * n,
		blockIdx.x * 256 + threadIdx.x);
}

// ---------------------------------------------------------------------------
// Background (GRmBackground2D.cu:11-94): out = in - (box mean of sqrt(in))^2
// over a periodic, Friedel-unfolded plane P(X, Y), X, Y in [0, T):
//   X <  T/2: P = in[Y][X];  X >= T/2: P = in[(T-Y) mod T][T-X].
// Box sums are separable: H = row sums, then column sums of H.
// ---------------------------------------------------------------------------
template <int T>
__global__ __launch_bounds__(256) void k_bg_rows(const float* __restrict__ raw, int log_spect,
	int half_box, float* __restrict__ hsum)
{
	constexpr int CX = T / 2 + 1;
	__shared__ float p[T];
	const float* in = raw + (long long)blockIdx.y * T * CX;
	float* out = hsum + (long long)blockIdx.y * T * CX;
	const int Y = blockIdx.x;
	for (int X = threadIdx.x; X < T; X += 256) {
		float v = X < T / 2 ? in[Y * CX + X] : in[((T - Y) & (T - 1)) * CX + (T - X)];
		p[X] = log_spect ? v : sqrtf(v);
	}
	__syncthreads();
	for (int x = threadIdx.x; x < CX; x += 256) {
		float acc = 0.f;
		for (int j = -half_box; j <= half_box; ++j) acc += p[(x + j + T) & (T - 1)];
		out[Y * CX + x] = acc;
	}
}

// Column sums, the subtraction, the r < 0.04 zeroing (evaluated in float as
// the reference does) and per-workgroup moment partials for the gate (an
// FP32 tree within the workgroup; k_gate adds the partials in FP64).
// Grid (ceil(T*CX/256), images).
template <int T>
__global__ __launch_bounds__(256) void k_bg_cols(const float* __restrict__ raw,
	const float* __restrict__ hsum, int log_spect, int half_box, float* __restrict__ spect,
	double* __restrict__ mom)
{
	constexpr int CX = T / 2 + 1;
	__shared__ float r1[256], r2[256];
	const int i = blockIdx.x * 256 + threadIdx.x;
	const float* in = raw + (long long)blockIdx.z * T * CX;
	const float* h = hsum + (long long)blockIdx.z * T * CX;
	float m1 = 0.0f, m2 = 0.0f;
	if (i < T * CX) {
		const int y = i / CX, x = i % CX;
		const int yc = y - T / 2;
		float fr = (float)x * 0.5f / (float)(CX - 1);
		fr = sqrtf(fr * fr + (float)(yc * yc) / (float)(T * T));
		float o;
		if ((double)fr < 0.04) {
			o = 0.f;
		} else {
			float acc = 0.f;
			for (int k = -half_box; k <= half_box; ++k) acc += h[((y + k + T) & (T - 1)) * CX + x];
			const int b = 2 * half_box + 1;
			acc /= (float)(b * b);
			if (!log_spect) acc *= acc;
			o = in[i] - acc;
		}
		spect[(long long)blockIdx.z * T * CX + i] = o;
		m1 = o; m2 = o * o;
	}
	r1[threadIdx.x] = m1; r2[threadIdx.x] = m2;
	__syncthreads();
	for (int s = 128; s > 0; s >>= 1) {
		if (threadIdx.x < s) { r1[threadIdx.x] += r1[threadIdx.x + s]; r2[threadIdx.x] += r2[threadIdx.x + s]; }
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		double* m = mom + ((long long)blockIdx.z * gridDim.x + blockIdx.x) * 2;
		m[0] = r1[0]; m[1] = r2[0];
	}
}

// Flatness gate (CFindCtfBase.cpp:167-174): low-pass iff var >= 1.
__global__ __launch_bounds__(256) void k_gate(const double* __restrict__ mom, int nparts, int n,
	SearchState* __restrict__ st)
{
	__shared__ double r1[256], r2[256];
	const double* m = mom + (long long)blockIdx.x * nparts * 2;
	double a = 0.0, b = 0.0;
	for (int i = threadIdx.x; i < nparts; i += 256) { a += m[2 * i]; b += m[2 * i + 1]; }
	r1[threadIdx.x] = a; r2[threadIdx.x] = b;
	__syncthreads();
	for (int s = 128; s > 0; s >>= 1) {
		if (threadIdx.x < s) { r1[threadIdx.x] += r1[threadIdx.x + s]; r2[threadIdx.x] += r2[threadIdx.x + s]; }
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		// CFindCtfBase.cpp:171-174: the two moments come back as floats and
		// fStd - fMean * fMean < 1 is evaluated in IEEE single (this library
		// is built without contraction, so the product is rounded first)
		const float fmean = (float)(r1[0] / n);
		float fstd = (float)(r2[0] / n);
		fstd = fstd - fmean * fmean;
		st[blockIdx.x].lowpassed = fstd < 1.0f ? 0 : 1;
	}
}

// ---------------------------------------------------------------------------
// Low-pass (CFindCtfBase.cpp:191-219, GLowpass2D.cu:9-23): the transfer
// exp(s (kx^2 + ky^2)) is separable, so the 2D FFT round trip equals a
// circular 1D round trip along x and then along y, each normalised by 1/T.
// x pass: rows of the unfolded full plane (GCalcSpectrum.cu:47-69, row T
// taken mod T); only the columns T/2 .. T (mod T) are kept.
// ---------------------------------------------------------------------------
template <int T>
__global__ __launch_bounds__(2 * (T / 8)) void k_lp_x(const float* __restrict__ spect,
	const SearchState* __restrict__ st, const float* __restrict__ h1,
	const float2* __restrict__ tw, float* __restrict__ q)
{
	constexpr int CX = T / 2 + 1, FT = T / 8, HX = T / 2;
	constexpr int FL = fft_lds_len(T);
	__shared__ float2 buf[2 * FL];
	if (!st[blockIdx.y].lowpassed) return;
	const int grp = threadIdx.x / FT, lane = threadIdx.x % FT;
	const int ya = 4 * blockIdx.x + 2 * grp;     // rows ya, ya + 1
	const float* in = spect + (long long)blockIdx.y * T * CX;
	float2* fb = buf + grp * FL;
	for (int xd = lane; xd < T; xd += FT) {
		float v[2];
#pragma unroll
		for (int r = 0; r < 2; ++r) {
			int yd = ya + r, xs = xd - HX, ys = yd;
			if (xs < 0) { xs = -xs; ys = (T - yd) & (T - 1); }
			v[r] = in[ys * CX + xs];
		}
		fb[pidx(xd)] = make_float2(v[0], v[1]);
	}
	fft_sync<FT>();
	fft_lds<T, FT, false>(fb, tw, lane);
	for (int k = lane; k < T; k += FT) {
		float f = h1[k];
		float2 z = fb[pidx(k)];
		fb[pidx(k)] = make_float2(z.x * f, z.y * f);
	}
	fft_sync<FT>();
	fft_lds<T, FT, true>(fb, tw, lane);
	float* out = q + (long long)blockIdx.y * T * CX;
	const float inv = 1.0f / (float)T;
	for (int c = lane; c < CX; c += FT) {
		float2 z = fb[pidx((HX + c) & (T - 1))];
		out[ya * CX + c] = z.x * inv;
		out[(ya + 1) * CX + c] = z.y * inv;
	}
}

// y pass on columns c, c+1 of q; result into spect.  Grid (ceil(CX/2), images).
template <int T>
__global__ __launch_bounds__(T / 8) void k_lp_y(const float* __restrict__ q,
	const SearchState* __restrict__ st, const float* __restrict__ h1,
	const float2* __restrict__ tw, float* __restrict__ spect)
{
	constexpr int CX = T / 2 + 1, FT = T / 8;
	__shared__ float2 buf[fft_lds_len(T)];
	if (!st[blockIdx.y].lowpassed) return;
	const int c0 = 2 * blockIdx.x;
	const bool two = c0 + 1 < CX;
	const float* in = q + (long long)blockIdx.y * T * CX;
	const int lane = threadIdx.x;
	for (int y = lane; y < T; y += FT)
		buf[pidx(y)] = make_float2(in[y * CX + c0], two ? in[y * CX + c0 + 1] : 0.f);
	__syncthreads();
	fft_lds<T, FT, false>(buf, tw, lane);
	for (int k = lane; k < T; k += FT) {
		float f = h1[k];
		float2 z = buf[pidx(k)];
		buf[pidx(k)] = make_float2(z.x * f, z.y * f);
	}
	__syncthreads();
	fft_lds<T, FT, true>(buf, tw, lane);
	float* out = spect + (long long)blockIdx.y * T * CX;
	const float inv = 1.0f / (float)T;
	for (int y = lane; y < T; y += FT) {
		float2 z = buf[pidx(y)];
		out[y * CX + c0] = z.x * inv;
		if (two) out[y * CX + c0 + 1] = z.y * inv;
	}
}

// ---------------------------------------------------------------------------
// Radial average (GRadialAvg.cu:11-33) from a host-built FP32 sample table.
// ---------------------------------------------------------------------------
// Grid (ceil(cmpx / 64), images), 64 threads: thread = bin, samples summed
// in table order in FP32.
__global__ __launch_bounds__(64) void k_radial(const float* __restrict__ spect, int cmpx, int cmpy,
	const int* __restrict__ roff, const int* __restrict__ rcnt, const int* __restrict__ ridx,
	float* __restrict__ radial, int ld)
{
	const float* in = spect + (long long)blockIdx.y * cmpx * cmpy;
	const int r = blockIdx.x * 64 + threadIdx.x;
	if (r >= cmpx) return;
	// FP32 in the reference's sample order (GRadialAvg.cu:24-31)
	float s = 0.0f;
	const int o = roff[r], n = rcnt[r];
	for (int i = 0; i < n; ++i) s += in[ridx[o + i]];
	radial[(long long)blockIdx.y * ld + r] = s / (float)n;
}

}  // namespace gctf
