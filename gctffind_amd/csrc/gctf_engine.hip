// gctf_engine.hip -- host engine and C-ABI of libgctf_hip.so.
//
// The engine owns per-shape plans (FFT tables, tile grid, ring tables),
// device buffers sized for a batch of micrographs, and the per-micrograph
// search state.  A batch runs as: [rescale -> tile spectrum] per chunk of
// images, then background -> gate -> low-pass -> radial average -> 1D
// search -> 2D search -> spectral CC for the whole batch.  No step waits on
// the host; the 2D search's control flow runs on the device, so a batch is
// captured into one HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gctf_hip.h"
#include "gctf_internal.hpp"
#include "k_spectrum.hpp"
#include "k_search.hpp"
#include "k_grid.hpp"
#include "gctf_net.hpp"

using namespace gctf;

namespace {

// ---------------------------------------------------------------------------
// small host helpers
// ---------------------------------------------------------------------------
struct DevBuf {
	void* p = nullptr;
	size_t bytes = 0;
	DevBuf() = default;
	DevBuf(const DevBuf&) = delete;
	DevBuf& operator=(const DevBuf&) = delete;
	~DevBuf() { release(); }
	void release()
	{
		if (p) (void)hipFree(p);
		p = nullptr;
		bytes = 0;
	}
	// grows to at least b bytes; returns true when the address changed
	hipError_t ensure(size_t b, bool* moved)
	{
		if (b <= bytes) return hipSuccess;
		release();
		hipError_t e = hipMalloc(&p, b);
		if (e != hipSuccess) { p = nullptr; return e; }
		bytes = b;
		if (moved) *moved = true;
		return hipSuccess;
	}
	template <class T> T* as() const { return static_cast<T*>(p); }
};

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
int next_pow2(int n)
{
	int m = 1;
	while (m < n) m <<= 1;
	return m;
}

void host_fft(std::vector<std::complex<double>>& a, bool inv)
{
	const int n = (int)a.size();
	for (int i = 1, j = 0; i < n; i++) {
		int bit = n >> 1;
		for (; j & bit; bit >>= 1) j ^= bit;
		j ^= bit;
		if (i < j) std::swap(a[i], a[j]);
	}
	for (int len = 2; len <= n; len <<= 1) {
		for (int k = 0; k < len / 2; k++) {
			double ang = (inv ? 2.0 : -2.0) * M_PI * k / len;
			std::complex<double> w(cos(ang), sin(ang));
			for (int i = k; i < n; i += len) {
				std::complex<double> u = a[i], v = a[i + len / 2] * w;
				a[i] = u + v;
				a[i + len / 2] = u - v;
			}
		}
	}
}

// Device tables of one transform length.
struct FftDev {
	int n = 0, m = 0;
	std::vector<unsigned char> rad;   // mixed-radix passes (empty: power of two or Bluestein)
	DevBuf tw, chirp, bhat;
	DevBuf sub_chirp, sub_bhat, sub_tw, ct_tw;   // radix-4 step over quarter-length Bluestein transforms
	Fft1dPlan plan() const
	{
		Fft1dPlan p{};
		p.n = n; p.m = m;
		p.tw = tw.as<float2>();
		p.chirp = chirp.p ? chirp.as<float2>() : nullptr;
		p.bhat = bhat.p ? bhat.as<float2>() : nullptr;
		p.nrad = (int)rad.size();
		for (size_t i = 0; i < rad.size(); ++i) p.rad[i] = rad[i];
		p.sub_chirp = sub_chirp.p ? sub_chirp.as<float2>() : nullptr;
		p.sub_bhat = sub_bhat.p ? sub_bhat.as<float2>() : nullptr;
		p.sub_tw = sub_tw.p ? sub_tw.as<float2>() : nullptr;
		p.ct_tw = ct_tw.p ? ct_tw.as<float2>() : nullptr;
		return p;
	}
};

// n = 2^a 3^b 5^c
bool smooth235(int n)
{
	if (n < 1) return false;
	while (n % 2 == 0) n /= 2;
	while (n % 3 == 0) n /= 3;
	while (n % 5 == 0) n /= 5;
	return n == 1;
}

// The length a transform of length n works on in LDS: n itself for powers
// of two and 2^a 3^b 5^c (Stockham passes), else the Bluestein length.
int work_len(int n) { return is_pow2(n) || smooth235(n) ? n : next_pow2(2 * n - 1); }

// Pass radices of a 2^a 3^b 5^c length: radix 8 while 8 divides what is
// left of the power of two, then one pass of 4 or 2, then the 3s and 5s.
std::vector<unsigned char> mixed_radices(int n)
{
	std::vector<unsigned char> r;
	int p2 = 0;
	while (n % 2 == 0) { n /= 2; ++p2; }
	for (; p2 >= 3; p2 -= 3) r.push_back(8);
	if (p2 == 2) r.push_back(4);
	if (p2 == 1) r.push_back(2);
	while (n % 3 == 0) { n /= 3; r.push_back(3); }
	while (n % 5 == 0) { n /= 5; r.push_back(5); }
	return r;
}

// Round-edge mask (GRoundEdge.cu:14-24) with centre (T/2, T/2), size (T, T).
double round_edge(int x, int y, int tile)
{
	double c = (double)((float)tile * 0.5f), s = (double)((float)tile * 1.0f);
	double fx = 2.0 * fabs((double)x - c) / s, fy = 2.0 * fabs((double)y - c) / s;
	double r = sqrt(fx * fx + fy * fy);
	if (r >= 1.0) return 0.0;
	double q = 0.5 * (1.0 - cos((double)3.1415926f * r));
	return 1.0 - q * q * q;
}

// Tables shared by every plan with tile size T.
struct TileTables {
	int T = 0;
	DevBuf tw, mask, h1, roff, rcnt, ridx, w1d;
	DevBuf boff, bcnt, bidx;   // pixels of each integer radius bin (spectral CC)
	int nsamples = 0;
	int nbins = 0;
};

struct RescalePlan {
	int nx = 0, ny = 0;
	float pix_raw = 0.f, bin = 1.f, pixn = 1.f;
	int nnx = 0, nny = 0, pitch = 0;
	bool active = false;
	// rows_half / c2r_half: the R2C / C2R row transforms run at half length
	// (even/odd samples packed into one complex row); rows / c2r are then the
	// half-length plans and rows_maxn / c2r_maxn the kernels' transform
	// lengths (8192 or 16384)
	bool rows_half = false, c2r_half = false;
	int rows_maxn = 16384, c2r_maxn = 16384;
	// cols_split: the column pass's forward transform runs as the transforms
	// of the even and odd samples; cf is then the half-length plan and
	// cols_maxn the kernel's transform length (8192 or 16384)
	bool cols_split = false;
	int cols_maxn = 16384;
	FftDev *rows = nullptr, *cf = nullptr, *ci = nullptr, *c2r = nullptr;
};

struct SearchPlan {
	float kv = 0, cs = 0, ac = 0, pix = 0;
	int T = 0;
	CtfConsts cc{};
	float lo1d = 0, hi1d = 0, flo2d = 0, fhi2d = 0;
	float res[2] = {0, 0};
	RingTables ring{};
	DevBuf idx, a, b, c2, s2, w;
};

struct GraphEntry {
	hipGraphExec_t exec = nullptr;
	int gen = -1;
};

// ---- static scratch of gctf_do2d_batch_static --------------------------------
// Search buffers for at most kStaticMaxImages half spectra of the 512 tile,
// in one file-scope device array: every GPU the library runs on holds its own
// copy until the process ends, at its own address (kStaticScratchBytes, about
// 18 MB).  Host code never passes `scratch` itself to a kernel (on the host
// it names a stand-in, not device memory); it uses the address that
// hipGetSymbolAddress returned for the device.
constexpr int kStaticMaxImages = 16;
constexpr int kStaticTile = 512;
constexpr int kStaticCx = kStaticTile / 2 + 1;
constexpr int kStaticMaxRing = 65536;   // ring pixels (zero-padded to 16) per spectrum
constexpr int kStaticMaxBins = 368;     // integer radius bins of the 512 half spectrum (363 used)
constexpr size_t st_align(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kStStates = 0;
constexpr size_t kStPlans = kStStates + st_align(sizeof(SearchState) * kStaticMaxImages);
constexpr size_t kStRefine = kStPlans + st_align(sizeof(ImgPlan) * kStaticMaxImages);
constexpr size_t kStRingSs = kStRefine + st_align(sizeof(RefineArgs) * kStaticMaxImages);
constexpr size_t kStRadial = kStRingSs + st_align(sizeof(double) * kStaticMaxImages);
constexpr size_t kStShellCc = kStRadial + st_align(sizeof(float) * kStaticCx * kStaticMaxImages);
constexpr size_t kStBins = kStShellCc + st_align(sizeof(float) * kStaticCx * kStaticMaxImages);
constexpr size_t kStPart = kStBins + st_align(sizeof(double) * 6 * kStaticMaxBins * kStaticMaxImages);
constexpr size_t kStRing = kStPart
	+ st_align(sizeof(double2) * (size_t)kMaxCands * kPixChunks * kStaticMaxImages);
constexpr size_t kStCtfMap = kStRing + st_align(sizeof(float) * (size_t)kStaticMaxRing * kStaticMaxImages);
// 1D search: the workgroup maxima of the reference's 501-point scan x 37 phases
constexpr int kStaticS1dLd = (501 * 37 + kS1dPerWg - 1) / kS1dPerWg;
constexpr size_t kStS1d = kStCtfMap + st_align(sizeof(float) * (size_t)kStaticTile * kStaticCx * kStaticMaxImages);
constexpr size_t kStaticScratchBytes = kStS1d + st_align(sizeof(float2) * (size_t)kStaticS1dLd * kStaticMaxImages);

}  // namespace

static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {

// Device address of `scratch` on each device, found once per device (with
// that device current) and kept until the process ends.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};

// The address of this device's scratch; `device` must be current.
hipError_t static_scratch_address(int device, void** out)
{
	*out = nullptr;
	if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
	std::lock_guard<std::mutex> lk(g_scratch_mu);
	if (!g_scratch_addr[device]) {
		void* p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess) return e;
		g_scratch_addr[device] = p;
	}
	*out = g_scratch_addr[device];
	return hipSuccess;
}

}  // namespace

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct gctf_ctx {
	int device = 0;
	hipStream_t own_stream = nullptr;
	hipStream_t stream = nullptr;
	gctf_params prm{};
	std::string err;
	int chunk = 4, tile_splits = 8, use_graph = 1;
	int cpb = 8;          // candidates per 2D-evaluator workgroup (search blocking: 4, 8 or 16)
	int list_queue = 1;   // tile row pass of a list: 1 persistent workgroups on a work queue, 0 one item each
	int tune_gen = 0;     // bumped by gctf_autotune (list graphs read the tuned values)
	float scan_lo = 0, scan_hi = 0, scan_step = 0;  // Do1D scan (gctf_set_search1d; 0: reference)
	int gen = 0;  // bumped whenever a device buffer moves

	std::map<int, std::unique_ptr<FftDev>> ffts;
	std::map<int, std::unique_ptr<TileTables>> tiles;
	std::map<std::string, std::unique_ptr<SearchPlan>> splans;
	std::map<std::string, GraphEntry> graphs;
	struct Tuned { int chunk, splits, cpb; };
	std::map<std::string, Tuned> tuned;  // shape -> gctf_autotune's choice

	// single-image (reference-style) state
	RescalePlan rp;
	bool rp_ready = false, rp_done = false;
	DevBuf pad1;
	DevBuf hostin;   // a host image's device copy for gctf_rescale, kept between calls
	// gctf_rescale_prefetch: the next host image's copy, made on copy_stream
	DevBuf prefetch;
	const float* prefetch_src = nullptr;
	size_t prefetch_bytes = 0;
	hipStream_t copy_stream = nullptr;
	hipEvent_t prefetch_done = nullptr, main_done = nullptr;
	float s_kv = 0, s_cs = 0, s_ac = 0, s_pix = 0;
	bool setup1 = false, setup2 = false, have_raw = false, have_spect = false;
	int s_nx = 0, s_ny = 0;

	// batch buffers
	int cap = 0;  // images the per-image buffers hold
	DevBuf cbufA, cbufB, padc, bsums, rowsum, rt, tpart;
	DevBuf s1part; // 1D search: workgroup maxima of a brute-force stage
	DevBuf jobs;   // tile work lists of gctf_estimate_list
	DevBuf rjobs, rtab, list_plans, qnext;   // its rescale jobs, class tables, plans, work-queue counters
	DevBuf raw, hsum, spect, mom, q, radial, ring_spec, ring_ss, states, part2d, refine, shell_cc, fullsp,
		moments1, plans, shell_bins;
	int kstride = 0;  // ring pixels reserved per image in ring_spec
	DevBuf g_bt, g_fn, g_gt, g_cc, g_part, g_top, g_rs, g_rescored;
	// learned path: proposal-network parameters, activations, outputs
	DevBuf net_params, net_out, props;
	DevBuf series;  // tilt-series controller state
	// device address of this GPU's static scratch (gctf_do2d_batch_static),
	// looked up at creation; null with static_err set when the lookup failed
	void* static_scratch = nullptr;
	std::string static_err;
	bool have_net = false;
	NetBuffers* netb = nullptr;
	std::vector<SearchState> h_states;
	SearchState* h_pinned = nullptr;
	int h_pinned_n = 0;

	// kernel timing with HIP events on `stream` (gctf_profile)
	struct ProfRec { const char* name; hipEvent_t a, b; };
	bool prof = false;
	std::vector<ProfRec> prec;
	std::vector<hipEvent_t> evpool;
	std::map<std::string, std::pair<double, int>> ptotal;

	~gctf_ctx()
	{
		for (auto& kv : graphs)
			if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
		if (h_pinned) (void)hipHostFree(h_pinned);
		net_buffers_destroy(netb);
		for (auto& r : prec) { evpool.push_back(r.a); evpool.push_back(r.b); }
		for (auto e : evpool) (void)hipEventDestroy(e);
		if (prefetch_done) (void)hipEventDestroy(prefetch_done);
		if (main_done) (void)hipEventDestroy(main_done);
		if (copy_stream) (void)hipStreamDestroy(copy_stream);
		if (own_stream) (void)hipStreamDestroy(own_stream);
	}
};

namespace {

int fail(gctf_ctx* c, int code, const std::string& msg)
{
	if (c) c->err = msg;
	return code;
}

#define HIPCHK(ctx, call)                                                                    \
	do {                                                                                     \
		hipError_t e_ = (call);                                                              \
		if (e_ != hipSuccess)                                                                \
			return fail(ctx, e_ == hipErrorOutOfMemory ? GCTF_ERR_NOMEM : GCTF_ERR_HIP,      \
				std::string(#call) + ": " + hipGetErrorString(e_));                           \
	} while (0)

#define CHK(expr)                    \
	do {                             \
		int r_ = (expr);             \
		if (r_ != GCTF_OK) return r_; \
	} while (0)

hipEvent_t prof_event(gctf_ctx* c)
{
	if (!c->evpool.empty()) { hipEvent_t e = c->evpool.back(); c->evpool.pop_back(); return e; }
	hipEvent_t e = nullptr;
	(void)hipEventCreate(&e);
	return e;
}

// Launch statement(s) bracketed by events on the context stream when
// profiling is on.
#define PROF(c, name, ...)                                                 \
	do {                                                                   \
		if ((c)->prof) {                                                   \
			hipEvent_t a_ = prof_event(c), b_ = prof_event(c);             \
			(void)hipEventRecord(a_, (c)->stream);                         \
			__VA_ARGS__;                                                   \
			(void)hipEventRecord(b_, (c)->stream);                         \
			(c)->prec.push_back({name, a_, b_});                           \
		} else {                                                           \
			__VA_ARGS__;                                                   \
		}                                                                  \
	} while (0)

int ensure_buf(gctf_ctx* c, DevBuf& b, size_t bytes)
{
	bool moved = false;
	hipError_t e = b.ensure(bytes, &moved);
	if (e != hipSuccess)
		return fail(c, GCTF_ERR_NOMEM, std::string("device allocation of ") + std::to_string(bytes)
			+ " bytes failed: " + hipGetErrorString(e));
	if (moved) c->gen++;
	return GCTF_OK;
}

int get_fft(gctf_ctx* c, int n, FftDev** out)
{
	auto it = c->ffts.find(n);
	if (it != c->ffts.end()) { *out = it->second.get(); return GCTF_OK; }
	auto f = std::make_unique<FftDev>();
	f->n = n;
	f->m = work_len(n);
	if (!is_pow2(n) && smooth235(n)) {
		f->rad = mixed_radices(n);
		if (f->rad.size() > sizeof(Fft1dPlan{}.rad)) return fail(c, GCTF_ERR_UNSUPPORTED, "transform length");
	}
	const int m = f->m;
	if (m > 16384)
		return fail(c, GCTF_ERR_UNSUPPORTED, "transform length " + std::to_string(n)
			+ " needs a working length above 16384");
	std::vector<float2> tw(m);
	for (int e = 0; e < m; ++e) {
		double a = -2.0 * M_PI * (double)e / (double)m;
		tw[e] = make_float2((float)cos(a), (float)sin(a));
	}
	CHK(ensure_buf(c, f->tw, sizeof(float2) * m));
	HIPCHK(c, hipMemcpy(f->tw.p, tw.data(), sizeof(float2) * m, hipMemcpyHostToDevice));
	// Bluestein tables of length n on m points: chirp[k] = exp(-pi i (k^2 mod
	// 2n) / n) and bhat = FFT_m of the wrapped conj(chirp)
	auto bluestein_tables = [&](int len, int work, DevBuf& dchirp, DevBuf& dbhat) -> int {
		std::vector<std::complex<double>> w(len), b(work, 0.0);
		std::vector<float2> chirp(len), bhat(work);
		for (int k = 0; k < len; ++k) {
			long long k2 = ((long long)k * k) % (2LL * len);
			double a = -M_PI * (double)k2 / (double)len;
			w[k] = std::complex<double>(cos(a), sin(a));
			chirp[k] = make_float2((float)w[k].real(), (float)w[k].imag());
		}
		b[0] = std::conj(w[0]);
		for (int k = 1; k < len; ++k) { b[k] = std::conj(w[k]); b[work - k] = b[k]; }
		host_fft(b, false);
		for (int k = 0; k < work; ++k) bhat[k] = make_float2((float)b[k].real(), (float)b[k].imag());
		CHK(ensure_buf(c, dchirp, sizeof(float2) * len));
		CHK(ensure_buf(c, dbhat, sizeof(float2) * work));
		HIPCHK(c, hipMemcpy(dchirp.p, chirp.data(), sizeof(float2) * len, hipMemcpyHostToDevice));
		HIPCHK(c, hipMemcpy(dbhat.p, bhat.data(), sizeof(float2) * work, hipMemcpyHostToDevice));
		return GCTF_OK;
	};
	auto twiddles = [&](int len, DevBuf& d) -> int {
		std::vector<float2> t(len);
		for (int e = 0; e < len; ++e) {
			double a = -2.0 * M_PI * (double)e / (double)len;
			t[e] = make_float2((float)cos(a), (float)sin(a));
		}
		CHK(ensure_buf(c, d, sizeof(float2) * len));
		HIPCHK(c, hipMemcpy(d.p, t.data(), sizeof(float2) * len, hipMemcpyHostToDevice));
		return GCTF_OK;
	};
	if (!is_pow2(n) && f->rad.empty()) {
		CHK(bluestein_tables(n, m, f->chirp, f->bhat));
		// n = 4 q whose quarter has Bluestein length m / 4: also the tables
		// of one radix-4 step over four q-point Bluestein transforms
		if (n % 4 == 0 && next_pow2(2 * (n / 4) - 1) == m / 4) {
			CHK(bluestein_tables(n / 4, m / 4, f->sub_chirp, f->sub_bhat));
			CHK(twiddles(m / 4, f->sub_tw));
			CHK(twiddles(n, f->ct_tw));
		}
	}
	*out = f.get();
	c->ffts[n] = std::move(f);
	return GCTF_OK;
}

// Radial sample table (GRadialAvg.cu:17-31) in IEEE single, the reference's
// operation order; the row index cmpy reached at r = cmpx-1 is taken mod cmpy.
void radial_table(int cmpx, int cmpy, std::vector<int>& off, std::vector<int>& cnt,
	std::vector<int>& idx)
{
	off.assign(cmpx, 0);
	cnt.assign(cmpx, 0);
	idx.clear();
	for (int r = 0; r < cmpx; ++r) {
		float fp = 3.1415926f * (float)r + 0.5f;
		int np = (int)fp;
		if (np < 1) np = 1;
		else if (np > 180) np = 180;
		float step = 180.0f / (float)np;
		off[r] = (int)idx.size();
		cnt[r] = np;
		for (int i = 0; i < np; ++i) {
			float rad = (-90.0f + (float)i * step) * 0.017453f;
			int x = (int)((float)r * cosf(rad) + 0.5f);
			int y = (int)((float)r * sinf(rad) + (float)(cmpy / 2) + 0.5f);
			idx.push_back((y % cmpy) * cmpx + x);
		}
	}
}

// Pixels of the half spectrum by integer radius bin floor(|(x, y - cmpy/2)|),
// the radius in the float arithmetic of the spectral-CC shells
// (GSpectralCC2D.cu:26-40).
void shell_bin_table(int cmpx, int cmpy, std::vector<int>& off, std::vector<int>& cnt, std::vector<int>& idx)
{
	std::vector<std::vector<int>> bins;
	for (int y = 0; y < cmpy; ++y) {
		const float fy = fabsf((float)y - (float)cmpy * 0.5f);
		for (int x = 0; x < cmpx; ++x) {
			const float fr = sqrtf((float)(x * x) + fy * fy);
			const int r = (int)fr;
			if (r >= (int)bins.size()) bins.resize(r + 1);
			bins[r].push_back(y * cmpx + x);
		}
	}
	off.assign(bins.size(), 0);
	cnt.assign(bins.size(), 0);
	idx.clear();
	for (size_t r = 0; r < bins.size(); ++r) {
		off[r] = (int)idx.size();
		cnt[r] = (int)bins[r].size();
		idx.insert(idx.end(), bins[r].begin(), bins[r].end());
	}
}

int get_tile_tables(gctf_ctx* c, int T, TileTables** out)
{
	auto it = c->tiles.find(T);
	if (it != c->tiles.end()) { *out = it->second.get(); return GCTF_OK; }
	auto t = std::make_unique<TileTables>();
	t->T = T;
	const int cx = T / 2 + 1;
	std::vector<float2> tw(T);
	for (int e = 0; e < T; ++e) {
		double a = -2.0 * M_PI * e / T;
		tw[e] = make_float2((float)cos(a), (float)sin(a));
	}
	std::vector<float> mask((size_t)T * T), h1(T);
	for (int y = 0; y < T; ++y)
		for (int x = 0; x < T; ++x) mask[(size_t)y * T + x] = (float)round_edge(x, y, T);
	// GLowpass2D::DoBFactor (GLowpass2D.cu:57-74): fScale = (float)(-10 / (Nx^2 + Ny^2))
	double dtemp = (double)T * T + (double)T * T;
	float fscale = (float)(-10.0f / dtemp);
	for (int k = 0; k < T; ++k) {
		int kk = k > T / 2 ? k - T : k;
		h1[k] = (float)exp((double)fscale * (double)(kk * kk));
	}
	std::vector<int> off, cnt, idx;
	radial_table(cx, T, off, cnt, idx);
	std::vector<int> boff, bcnt, bidx;
	shell_bin_table(cx, T, boff, bcnt, bidx);
	t->nbins = (int)boff.size();
	CHK(ensure_buf(c, t->boff, sizeof(int) * boff.size()));
	CHK(ensure_buf(c, t->bcnt, sizeof(int) * bcnt.size()));
	CHK(ensure_buf(c, t->bidx, sizeof(int) * bidx.size()));
	HIPCHK(c, hipMemcpy(t->boff.p, boff.data(), sizeof(int) * boff.size(), hipMemcpyHostToDevice));
	HIPCHK(c, hipMemcpy(t->bcnt.p, bcnt.data(), sizeof(int) * bcnt.size(), hipMemcpyHostToDevice));
	HIPCHK(c, hipMemcpy(t->bidx.p, bidx.data(), sizeof(int) * bidx.size(), hipMemcpyHostToDevice));
	std::vector<double> w1d(cx);
	for (int x = 0; x < cx; ++x) {
		double f = (double)x / ((double)cx - 1.0);  // GCC1D.cu:36-39, B = 1
		w1d[x] = exp(-f * f);
	}
	CHK(ensure_buf(c, t->tw, sizeof(float2) * T));
	CHK(ensure_buf(c, t->mask, sizeof(float) * mask.size()));
	CHK(ensure_buf(c, t->h1, sizeof(float) * T));
	CHK(ensure_buf(c, t->roff, sizeof(int) * cx));
	CHK(ensure_buf(c, t->rcnt, sizeof(int) * cx));
	CHK(ensure_buf(c, t->ridx, sizeof(int) * idx.size()));
	CHK(ensure_buf(c, t->w1d, sizeof(double) * cx));
	HIPCHK(c, hipMemcpy(t->tw.p, tw.data(), sizeof(float2) * T, hipMemcpyHostToDevice));
	HIPCHK(c, hipMemcpy(t->mask.p, mask.data(), sizeof(float) * mask.size(), hipMemcpyHostToDevice));
	HIPCHK(c, hipMemcpy(t->h1.p, h1.data(), sizeof(float) * T, hipMemcpyHostToDevice));
	HIPCHK(c, hipMemcpy(t->roff.p, off.data(), sizeof(int) * cx, hipMemcpyHostToDevice));
	HIPCHK(c, hipMemcpy(t->rcnt.p, cnt.data(), sizeof(int) * cx, hipMemcpyHostToDevice));
	HIPCHK(c, hipMemcpy(t->ridx.p, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice));
	HIPCHK(c, hipMemcpy(t->w1d.p, w1d.data(), sizeof(double) * cx, hipMemcpyHostToDevice));
	t->nsamples = (int)idx.size();
	*out = t.get();
	c->tiles[T] = std::move(t);
	return GCTF_OK;
}

// CCTFTheory::Setup (CCTFTheory.cpp:69-86) + the kernels' amplitude phase
// (GCalcCTF2D.cu:88-97) + resolution range (CFindCtfBase.cpp:64-68) + CC
// frequency bounds (CFindDefocus1D.cpp:128-135, CFindDefocus2D.cpp:86-92).
std::string plan_key(float kv, float cs, float ac, float pix, int T, float res_lo, float res_hi)
{
	char key[160];
	snprintf(key, sizeof key, "%a/%a/%a/%a/%d/%a/%a", kv, cs, ac, pix, T, res_lo, res_hi);
	return key;
}

// An existing plan with the default ring, or null; creates nothing.
SearchPlan* find_search_plan(gctf_ctx* c, float kv, float cs, float ac, float pix, int T)
{
	auto it = c->splans.find(plan_key(kv, cs, ac, pix, T, 0.f, 0.f));
	return it == c->splans.end() ? nullptr : it->second.get();
}

int get_search_plan(gctf_ctx* c, float kv, float cs, float ac, float pix, int T, SearchPlan** out,
	float res_lo = 0.f, float res_hi = 0.f)
{
	const std::string key = plan_key(kv, cs, ac, pix, T, res_lo, res_hi);
	auto it = c->splans.find(key);
	if (it != c->splans.end()) { *out = it->second.get(); return GCTF_OK; }
	auto sp = std::make_unique<SearchPlan>();
	sp->kv = kv; sp->cs = cs; sp->ac = ac; sp->pix = pix; sp->T = T;
	float kv1000 = kv * 1000.0f;
	float wl_a = (float)(12.26 / sqrt((double)kv1000 + 0.9784 * (double)kv * (double)kv));
	sp->cc.wl = wl_a / pix;
	sp->cc.cs = (float)((double)cs * 1e7 / (double)pix);
	sp->cc.amp_phase = atanf(ac / sqrtf(1.0f - ac * ac));
	sp->cc.pix = pix;
	sp->res[0] = 25.0f * pix;
	sp->res[1] = (2.0f * pix) / 0.8f;
	if (sp->res[1] < 3.5f) sp->res[1] = 3.5f;
	if (res_lo > 0.f && res_hi > 0.f) { sp->res[0] = res_lo; sp->res[1] = res_hi; }
	const int cx = T / 2 + 1, cy = T;
	float r1 = (float)((cx - 1) * 2) * pix;
	sp->lo1d = r1 / sp->res[0];
	sp->hi1d = r1 / sp->res[1];
	float r2 = (float)cy * pix;
	sp->flo2d = r2 / sp->res[0];
	sp->fhi2d = r2 / sp->res[1];
	// ring (GCC2D.cu:28-33,152-156), float membership test
	float l = sp->flo2d / (float)cy, h = sp->fhi2d / (float)cy;
	if (h > 0.75f) h = 0.75f;
	float lo2 = l * l, hi2 = h * h;
	std::vector<int> idx;
	std::vector<float> a, b, c2, s2, w;
	const double pir = (double)3.1415926f / M_PI;
	const double wl = sp->cc.wl, csp = sp->cc.cs;
	for (int y = 0; y < cy; ++y)
		for (int x = 0; x < cx; ++x) {
			float fy = ((float)y - (float)cy * 0.5f) / (float)cy;
			float fx = (0.5f * (float)x) / (float)(cx - 1);
			float s2f = fx * fx + fy * fy;
			if (s2f < lo2 || s2f > hi2) continue;
			double dfx = (double)((float)x * 0.5f / (float)(cx - 1));
			double dfy = (double)((float)(y - cy / 2) / (float)cy);
			double ss = dfx * dfx + dfy * dfy;
			double th = atan(dfy / (dfx + (double)1e-30f));
			idx.push_back(y * cx + x);
			a.push_back((float)(pir * wl * ss));
			b.push_back((float)(pir * 0.5 * wl * wl * wl * ss * ss * csp));
			c2.push_back((float)cos(2.0 * th));
			s2.push_back((float)sin(2.0 * th));
			w.push_back((float)exp(-16.0 * (double)s2f));
		}
	const int n = (int)idx.size();
	if (n == 0) return fail(c, GCTF_ERR_ARG, "empty correlation ring");
	// zero-padded to a multiple of 16 (w = 0 there) for the grid GEMM's vector loads
	const int n16 = (n + 15) / 16 * 16;
	idx.resize(n16, 0); a.resize(n16, 0.f); b.resize(n16, 0.f); c2.resize(n16, 0.f); s2.resize(n16, 0.f);
	w.resize(n16, 0.f);
	CHK(ensure_buf(c, sp->idx, sizeof(int) * n16));
	CHK(ensure_buf(c, sp->a, sizeof(float) * n16));
	CHK(ensure_buf(c, sp->b, sizeof(float) * n16));
	CHK(ensure_buf(c, sp->c2, sizeof(float) * n16));
	CHK(ensure_buf(c, sp->s2, sizeof(float) * n16));
	CHK(ensure_buf(c, sp->w, sizeof(float) * n16));
	HIPCHK(c, hipMemcpy(sp->idx.p, idx.data(), sizeof(int) * n16, hipMemcpyHostToDevice));
	HIPCHK(c, hipMemcpy(sp->a.p, a.data(), sizeof(float) * n16, hipMemcpyHostToDevice));
	HIPCHK(c, hipMemcpy(sp->b.p, b.data(), sizeof(float) * n16, hipMemcpyHostToDevice));
	HIPCHK(c, hipMemcpy(sp->c2.p, c2.data(), sizeof(float) * n16, hipMemcpyHostToDevice));
	HIPCHK(c, hipMemcpy(sp->s2.p, s2.data(), sizeof(float) * n16, hipMemcpyHostToDevice));
	HIPCHK(c, hipMemcpy(sp->w.p, w.data(), sizeof(float) * n16, hipMemcpyHostToDevice));
	sp->ring.n = n;
	sp->ring.idx = sp->idx.as<int>();
	sp->ring.a = sp->a.as<float>();
	sp->ring.b = sp->b.as<float>();
	sp->ring.c2 = sp->c2.as<float>();
	sp->ring.s2 = sp->s2.as<float>();
	sp->ring.w = sp->w.as<float>();
	*out = sp.get();
	c->splans[key] = std::move(sp);
	return GCTF_OK;
}

// CRescaleImage::Setup sizes (CRescaleImage.cpp:47-73).
int make_rescale_plan(gctf_ctx* c, int nx, int ny, float pix, RescalePlan& rp)
{
	if (nx < 2 || ny < 2 || !(pix > 0)) return fail(c, GCTF_ERR_ARG, "bad image size or pixel size");
	rp.nx = nx; rp.ny = ny; rp.pix_raw = pix;
	float bin = 1.2f / pix;
	if (bin <= 1) bin = 1.0f;
	rp.bin = bin;
	rp.pixn = pix * bin;
	int nnx = (int)((float)nx / bin + 0.5f), nny = (int)((float)ny / bin + 0.5f);
	rp.nnx = nnx / 2 * 2;
	rp.nny = nny / 2 * 2;
	rp.pitch = (rp.nnx / 2 + 1) * 2;
	rp.active = bin > 1;
	if (rp.active) {
		// an even row length whose working length exceeds one LDS transform
		// runs as a half-length complex transform
		auto half = [](int n) { return n % 2 == 0 && !is_pow2(n) && work_len(n) > 16384; };
		rp.rows_half = half(nx);
		rp.c2r_half = half(rp.nnx);
		// a row whose working length is above 8192 also runs at half length
		// when that fits 8192-point transforms: 73.7 KB of LDS, two
		// workgroups per CU
		auto half8 = [](int n) {
			return n % 2 == 0 && !is_pow2(n) && work_len(n) > 8192 && work_len(n) <= 16384 && work_len(n / 2) <= 8192;
		};
		if (half8(nx)) { rp.rows_half = true; rp.rows_maxn = 8192; }
		if (half8(rp.nnx)) { rp.c2r_half = true; rp.c2r_maxn = 8192; }
		rp.cols_split = half(ny);
		if (rp.cols_split && rp.nny > 8192)
			return fail(c, GCTF_ERR_UNSUPPORTED, "rescaled height " + std::to_string(rp.nny)
				+ " above 8192 for a split column transform");
		// a column whose working length is above 8192 also splits when its
		// halves fit 8192-point transforms and the rescaled height is at most
		// 4096 (the split kernel keeps up to MAXN/2 rows): 73.7 KB of LDS, two
		// workgroups per CU instead of one
		if (!rp.cols_split && ny % 2 == 0 && !is_pow2(ny) && work_len(ny) > 8192 && work_len(ny) <= 16384
			&& work_len(ny / 2) <= 8192 && rp.nny <= 4096) {
			rp.cols_split = true;
			rp.cols_maxn = 8192;
		}
		CHK(get_fft(c, rp.rows_half ? nx / 2 : nx, &rp.rows));
		CHK(get_fft(c, rp.cols_split ? ny / 2 : ny, &rp.cf));
		CHK(get_fft(c, rp.nny, &rp.ci));
		CHK(get_fft(c, rp.c2r_half ? rp.nnx / 2 : rp.nnx, &rp.c2r));
	}
	return GCTF_OK;
}

TileGrid make_grid(int nx, int ny, int T)
{
	TileGrid g{};
	int ov = (int)((float)T * 0.50f);
	ov = ov / 2 * 2;
	int step = T - ov;
	g.tile = T;
	g.step = step;
	g.ntx = (nx - ov) / step;
	g.nty = (ny - ov) / step;
	g.offx = (nx - g.ntx * T + (g.ntx - 1) * ov) / 2;
	g.offy = (ny - g.nty * T + (g.nty - 1) * ov) / 2;
	g.blocks = (2 * step == T) ? 1 : 0;
	if (g.blocks) { g.nbx = g.ntx + 1; g.nby = g.nty + 1; g.bsize = step; }
	else { g.nbx = g.ntx; g.nby = g.nty; g.bsize = T; }
	return g;
}

int check_tile(gctf_ctx* c, int T)
{
	if (T != 256 && T != 512 && T != 1024)
		return fail(c, GCTF_ERR_UNSUPPORTED, "tile size " + std::to_string(T)
			+ " not supported (256, 512 or 1024)");
	return GCTF_OK;
}

// Workgroup maxima per image of a 1D brute-force stage: the largest stage
// is the initial scan (the reference's 501 points, or the configured scan)
// times 37 phases, or the refine stage (at most 501 x 37).
// Workgroups of a 1D stage with at most nps phases.
int s1d_groups(const gctf_ctx* c, int nps)
{
	long long ndf = 501;
	if (c->scan_step > 0.0f)
		ndf = std::max(ndf, (long long)((double)(c->scan_hi - c->scan_lo) / (double)c->scan_step) + 3);
	return (int)((ndf * nps + kS1dPerWg - 1) / kS1dPerWg);
}
int s1d_part_ld(const gctf_ctx* c) { return s1d_groups(c, 37); }

// Per-image buffers for `n` images of tile size T.
int ensure_image_buffers(gctf_ctx* c, int n, int T, int ring_n)
{
	CHK(ensure_buf(c, c->s1part, sizeof(float2) * (size_t)s1d_part_ld(c) * n));
	const size_t cx = T / 2 + 1, npix = cx * T;
	const int nparts = (int)((npix + 255) / 256);
	CHK(ensure_buf(c, c->raw, sizeof(float) * npix * n));
	CHK(ensure_buf(c, c->hsum, sizeof(float) * npix * n));
	CHK(ensure_buf(c, c->spect, sizeof(float) * npix * n));
	CHK(ensure_buf(c, c->q, sizeof(float) * npix * n));
	CHK(ensure_buf(c, c->mom, sizeof(double) * 2 * nparts * n));
	CHK(ensure_buf(c, c->radial, sizeof(float) * cx * n));
	CHK(ensure_buf(c, c->ring_spec, sizeof(float) * (size_t)ring_n * n));
	CHK(ensure_buf(c, c->ring_ss, sizeof(double) * n));
	CHK(ensure_buf(c, c->part2d, sizeof(double2) * (size_t)kMaxCands * kPixChunks * n));
	CHK(ensure_buf(c, c->refine, sizeof(RefineArgs) * n));
	CHK(ensure_buf(c, c->plans, sizeof(ImgPlan) * n));
	CHK(ensure_buf(c, c->shell_cc, sizeof(float) * cx * n));
	// integer radius bins of the half spectrum: floor(sqrt(cx^2 + (T/2)^2)) + 1 at most
	const size_t nb = (size_t)std::sqrt((double)cx * cx + (double)(T / 2) * (T / 2)) + 2;
	CHK(ensure_buf(c, c->shell_bins, sizeof(double) * 6 * nb * n));
	size_t old = c->states.bytes;
	CHK(ensure_buf(c, c->states, sizeof(SearchState) * n));
	if (c->states.bytes != old) {
		// new state storage starts zeroed (CFindCtfBase / CFindDefocus2D constructors)
		std::vector<SearchState> z(c->states.bytes / sizeof(SearchState));
		memset(z.data(), 0, sizeof(SearchState) * z.size());
		HIPCHK(c, hipMemcpy(c->states.p, z.data(), sizeof(SearchState) * z.size(), hipMemcpyHostToDevice));
	}
	if (c->h_pinned_n < n) {
		if (c->h_pinned) (void)hipHostFree(c->h_pinned);
		c->h_pinned = nullptr;
		HIPCHK(c, hipHostMalloc((void**)&c->h_pinned, sizeof(SearchState) * n, hipHostMallocDefault));
		c->h_pinned_n = n;
	}
	return GCTF_OK;
}

// ---------------------------------------------------------------------------
// launch sequences
// ---------------------------------------------------------------------------

// Complex elements of one micrograph's part of each rescale buffer: the row
// spectra (ny rows, spec_ld(ncx) apart) or the rescaled columns (nny x ncx),
// rounded up to 256 bytes so that every micrograph's rows start on a line.
size_t cbuf_elems(const RescalePlan& rp)
{
	const int ncx = rp.nnx / 2 + 1;
	const size_t e = std::max((size_t)rp.ny * spec_ld(ncx), (size_t)ncx * rp.nny);
	return (e + 31) / 32 * 32;
}

// Rescale n images (CRescaleImage::DoIt) into padded rows at out + i * out_stride.
// With tg, the C2R pass also leaves the row partials of tg's mean blocks in
// c->rowsum (launch_spectrum then reads them instead of the image).
template <typename Tin>
int launch_rescale_t(gctf_ctx* c, const RescalePlan& rp, const Tin* in, long long in_stride, int n,
	float* out, long long out_stride, const TileGrid* tg, float* rowsum_out)
{
	hipStream_t s = c->stream;
	const int ncx = rp.nnx / 2 + 1;
	float2* A = c->cbufA.as<float2>();
	float2* B = c->cbufB.as<float2>();
	const long long stA = (long long)cbuf_elems(rp), stB = stA;
	// 1. rows R2C with x crop into A: A[y][kx], rows spec_ld(ncx) apart
	{
		const int m = rp.rows->m;
		dim3 grid((rp.ny + 1) / 2, n);
		Fft1dPlan p = rp.rows->plan();
		if (rp.rows_half) {
			if (m > rp.rows_maxn) return fail(c, GCTF_ERR_UNSUPPORTED, "row length");
			if (rp.rows_maxn == 8192)
				PROF(c, "k_rescale_rows_r2c_half", k_rescale_rows_r2c_half<Tin, 1024, 8192><<<dim3(rp.ny, n), 1024,
					0, s>>>(in, in_stride, rp.nx, rp.ny, p, ncx, A, stA));
			else
				PROF(c, "k_rescale_rows_r2c_half", k_rescale_rows_r2c_half<Tin, 1024, 16384><<<dim3(rp.ny, n), 1024,
					0, s>>>(in, in_stride, rp.nx, rp.ny, p, ncx, A, stA));
		} else
		PROF(c, "k_rescale_rows_r2c", if (m <= 2048) k_rescale_rows_r2c<Tin, 256, 2048><<<grid, 256, 0, s>>>(in, in_stride, rp.nx, rp.ny, p, ncx, A, stA);
			else if (m <= 4096) k_rescale_rows_r2c<Tin, 512, 4096><<<grid, 512, 0, s>>>(in, in_stride, rp.nx, rp.ny, p, ncx, A, stA);
			else if (m <= 8192) k_rescale_rows_r2c<Tin, 1024, 8192><<<grid, 1024, 0, s>>>(in, in_stride, rp.nx, rp.ny, p, ncx, A, stA);
			else k_rescale_rows_r2c<Tin, 1024, 16384><<<grid, 1024, 0, s>>>(in, in_stride, rp.nx, rp.ny, p, ncx, A, stA));
	}
	// transpose (ny x ncx) -> B[kx][y]
	{
		dim3 grid((ncx + 31) / 32, (rp.ny + 31) / 32, n);
		PROF(c, "k_transpose_c", k_transpose_c<<<grid, 256, 0, s>>>(A, stA, rp.ny, ncx, B, stB));
	}
	// 2. columns: forward ny, crop, inverse nny, written row-major into A: A[y'][kx]
	{
		const int m = std::max(rp.cf->m, rp.ci->m);
		dim3 grid(8 * ((ncx + 7) / 8), n);
		Fft1dPlan pf = rp.cf->plan(), pi = rp.ci->plan();
		if (rp.cols_split) {
			if (m > rp.cols_maxn) return fail(c, GCTF_ERR_UNSUPPORTED, "column length");
			if (rp.cols_maxn == 8192)
				PROF(c, "k_rescale_cols_split", k_rescale_cols_split<1024, 8192><<<grid, 1024, 0, s>>>(B, stB, rp.ny,
					rp.nny, ncx, pf, pi, A, stA));
			else
				PROF(c, "k_rescale_cols_split", k_rescale_cols_split<1024, 16384><<<grid, 1024, 0, s>>>(B, stB, rp.ny,
					rp.nny, ncx, pf, pi, A, stA));
		} else
		PROF(c, "k_rescale_cols", if (m <= 2048) k_rescale_cols<256, 2048><<<grid, 256, 0, s>>>(B, stB, rp.ny, rp.nny, ncx, pf, pi, A, stA);
			else if (m <= 4096) k_rescale_cols<512, 4096><<<grid, 512, 0, s>>>(B, stB, rp.ny, rp.nny, ncx, pf, pi, A, stA);
			else if (m <= 8192) k_rescale_cols<1024, 8192><<<grid, 1024, 0, s>>>(B, stB, rp.ny, rp.nny, ncx, pf, pi, A, stA);
			else k_rescale_cols<1024, 16384><<<grid, 1024, 0, s>>>(B, stB, rp.ny, rp.nny, ncx, pf, pi, A, stA));
	}
	// 3. rows C2R into the padded image
	{
		const int m = rp.c2r->m;
		dim3 grid((rp.nny + 1) / 2, n);
		Fft1dPlan p = rp.c2r->plan();
		const TileGrid g = tg ? *tg : TileGrid{};
		float* rs = tg ? (rowsum_out ? rowsum_out : c->rowsum.as<float>()) : nullptr;
		if (rp.c2r_half) {
			if (m > rp.c2r_maxn) return fail(c, GCTF_ERR_UNSUPPORTED, "row length");
			if (rp.c2r_maxn == 8192)
				PROF(c, "k_rescale_rows_c2r_half", k_rescale_rows_c2r_half<1024, 8192><<<dim3(rp.nny, n), 1024, 0,
					s>>>(A, stA, rp.nny, rp.nnx, p, out, out_stride, rp.pitch, g, rs));
			else
				PROF(c, "k_rescale_rows_c2r_half", k_rescale_rows_c2r_half<1024, 16384><<<dim3(rp.nny, n), 1024, 0,
					s>>>(A, stA, rp.nny, rp.nnx, p, out, out_stride, rp.pitch, g, rs));
		} else
		PROF(c, "k_rescale_rows_c2r", if (m <= 2048) k_rescale_rows_c2r<256, 2048><<<grid, 256, 0, s>>>(A, stA, rp.nny, rp.nnx, p, out, out_stride, rp.pitch, g, rs);
			else if (m <= 4096) k_rescale_rows_c2r<512, 4096><<<grid, 512, 0, s>>>(A, stA, rp.nny, rp.nnx, p, out, out_stride, rp.pitch, g, rs);
			else if (m <= 8192) k_rescale_rows_c2r<1024, 8192><<<grid, 1024, 0, s>>>(A, stA, rp.nny, rp.nnx, p, out, out_stride, rp.pitch, g, rs);
			else k_rescale_rows_c2r<1024, 16384><<<grid, 1024, 0, s>>>(A, stA, rp.nny, rp.nnx, p, out, out_stride, rp.pitch, g, rs));
	}
	HIPCHK(c, hipGetLastError());
	return GCTF_OK;
}

int ensure_rescale_buffers(gctf_ctx* c, const RescalePlan& rp, int n)
{
	const size_t per = cbuf_elems(rp);
	CHK(ensure_buf(c, c->cbufA, sizeof(float2) * per * n));
	CHK(ensure_buf(c, c->cbufB, sizeof(float2) * per * n));
	return GCTF_OK;
}

int launch_rescale(gctf_ctx* c, const RescalePlan& rp, const void* in, int dtype, long long in_stride,
	int n, float* out, long long out_stride, const TileGrid* tg = nullptr, float* rowsum_out = nullptr)
{
	if (dtype == GCTF_DTYPE_F16)
		return launch_rescale_t<_Float16>(c, rp, (const _Float16*)in, in_stride, n, out, out_stride, tg,
			rowsum_out);
	return launch_rescale_t<float>(c, rp, (const float*)in, in_stride, n, out, out_stride, tg, rowsum_out);
}

int ensure_spectrum_buffers(gctf_ctx* c, const TileGrid& g, int T, int n)
{
	const size_t cx = T / 2 + 1;
	const size_t ntiles = (size_t)g.ntx * g.nty;
	CHK(ensure_buf(c, c->bsums, sizeof(float) * (size_t)g.nbx * g.nby * kSumParts * n));
	const size_t nr = (size_t)(g.nby - 1) * g.step + g.bsize;   // rows of the mean blocks
	CHK(ensure_buf(c, c->rowsum, sizeof(float) * nr * g.nbx * n));
	CHK(ensure_buf(c, c->rt, sizeof(float2) * ntiles * cx * T * n));
	CHK(ensure_buf(c, c->tpart, sizeof(float) * (size_t)c->tile_splits * T * cx * n));
	return GCTF_OK;
}

// Tile-averaged amplitude spectra (CGenAvgSpectrum::DoIt) of n images into
// raw[slot0 .. slot0+n).  rows_ready: the rescale left the mean blocks' row
// partials in c->rowsum.
template <typename Tin, int T>
int launch_spectrum_t(gctf_ctx* c, const TileGrid& g, const TileTables& tt, const Tin* img, int pitch,
	long long stride, int n, int slot0, bool rows_ready)
{
	hipStream_t s = c->stream;
	constexpr int CX = T / 2 + 1;
	const int ntiles = g.ntx * g.nty;
	const int nsplit = std::min(c->tile_splits, ntiles);
	if (!rows_ready) {
		const int nr = (g.nby - 1) * g.step + g.bsize;
		PROF(c, "k_row_sums", k_row_sums<Tin><<<dim3(nr, n), 256, 0, s>>>(img, stride, pitch, g,
			c->rowsum.as<float>()));
	}
	const int np = g.nbx * g.nby * kSumParts;
	PROF(c, "k_block_rows", k_block_rows<<<dim3((np + 255) / 256, n), 256, 0, s>>>(c->rowsum.as<float>(), g,
		c->bsums.as<float>()));
	constexpr int ROWS = 16;                                 // rows per row-pass workgroup
	const long long rts = (long long)ntiles * CX * T;
	constexpr int RG = 2;                                    // row blocks per row-pass workgroup
	PROF(c, "k_tile_rows", k_tile_rows<Tin, T, ROWS, RG><<<dim3((ntiles * (T / ROWS) + RG - 1) / RG, n),
		(ROWS / 2) * (T / 8), 0, s>>>(
		img, stride, pitch, g, c->bsums.as<float>(), tt.mask.as<float>(), tt.tw.as<float2>(),
		c->rt.as<float2>(), rts));
	constexpr int CPW = 4;
	const float fac = 1.0f / (float)ntiles;                   // CGenAvgSpectrum.cpp:90
	PROF(c, "k_tile_cols", k_tile_cols<T, CPW><<<dim3((CX + CPW - 1) / CPW, nsplit, n), CPW * (T / 8), 0, s>>>(
		c->rt.as<float2>(), rts, ntiles, nsplit, fac, tt.tw.as<float2>(), c->tpart.as<float>()));
	const int npix = T * CX;
	PROF(c, "k_tile_reduce", k_tile_reduce<<<dim3((npix + 255) / 256, n), 256, 0, s>>>(c->tpart.as<float>(), nsplit, npix,
		c->raw.as<float>() + (size_t)slot0 * npix));
	HIPCHK(c, hipGetLastError());
	return GCTF_OK;
}

template <typename Tin>
int launch_spectrum(gctf_ctx* c, const TileGrid& g, const TileTables& tt, const Tin* img, int pitch,
	long long stride, int n, int slot0, bool rows_ready = false)
{
	switch (g.tile) {
	case 256: return launch_spectrum_t<Tin, 256>(c, g, tt, img, pitch, stride, n, slot0, rows_ready);
	case 512: return launch_spectrum_t<Tin, 512>(c, g, tt, img, pitch, stride, n, slot0, rows_ready);
	case 1024: return launch_spectrum_t<Tin, 1024>(c, g, tt, img, pitch, stride, n, slot0, rows_ready);
	default: return fail(c, GCTF_ERR_UNSUPPORTED, "tile size");
	}
}

// Background removal, flatness gate and low-pass (CFindCtfBase::
// mRemoveBackground / mLowpass) for images [0, n).
template <int T>
int launch_post_t(gctf_ctx* c, const TileTables& tt, int n)
{
	hipStream_t s = c->stream;
	constexpr int CX = T / 2 + 1;
	const int npix = T * CX;
	// box side (GRmBackground2D.cu:33-36) with fMinFreq = 1/15
	int box = (int)((1.0f / 15.0f) * (float)T);
	if ((float)box > 0.3f * (float)T) box = (int)(0.3f * (float)T);
	box = box / 2 * 2 + 1;
	if (box < 7) box = 7;
	const int hb = box / 2;
	const int ls = c->prm.log_spect ? 1 : 0;
	PROF(c, "k_bg_rows", k_bg_rows<T><<<dim3(T, n), 256, 0, s>>>(c->raw.as<float>(), ls, hb, c->hsum.as<float>()));
	const int nparts = (npix + 255) / 256;
	PROF(c, "k_bg_cols", k_bg_cols<T><<<dim3(nparts, 1, n), 256, 0, s>>>(c->raw.as<float>(), c->hsum.as<float>(), ls, hb,
		c->spect.as<float>(), c->mom.as<double>()));
	PROF(c, "k_gate", k_gate<<<n, 256, 0, s>>>(c->mom.as<double>(), nparts, npix, c->states.as<SearchState>()));
	PROF(c, "k_lp_x", k_lp_x<T><<<dim3(T / 4, n), 2 * (T / 8), 0, s>>>(c->spect.as<float>(), c->states.as<SearchState>(),
		tt.h1.as<float>(), tt.tw.as<float2>(), c->q.as<float>()));
	PROF(c, "k_lp_y", k_lp_y<T><<<dim3((CX + 1) / 2, n), T / 8, 0, s>>>(c->q.as<float>(), c->states.as<SearchState>(),
		tt.h1.as<float>(), tt.tw.as<float2>(), c->spect.as<float>()));
	HIPCHK(c, hipGetLastError());
	return GCTF_OK;
}

int launch_post(gctf_ctx* c, const TileTables& tt, int n)
{
	switch (tt.T) {
	case 256: return launch_post_t<256>(c, tt, n);
	case 512: return launch_post_t<512>(c, tt, n);
	case 1024: return launch_post_t<1024>(c, tt, n);
	default: return fail(c, GCTF_ERR_UNSUPPORTED, "tile size");
	}
}

// Per-image search plans: images [i0, i0 + m) get the plan of `sp`.
__global__ void k_fill_plans(ImgPlan* __restrict__ dst, int m, ImgPlan v)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < m) dst[i] = v;
}

// The plan of `sp` into dst[0 .. m).
int fill_plans_at(gctf_ctx* c, ImgPlan* dst, int m, const SearchPlan& sp)
{
	ImgPlan v;
	v.cc = sp.cc;
	v.ring = sp.ring;
	v.lo1d = sp.lo1d;
	v.hi1d = sp.hi1d;
	k_fill_plans<<<(m + 63) / 64, 64, 0, c->stream>>>(dst, m, v);
	HIPCHK(c, hipGetLastError());
	return GCTF_OK;
}

int fill_plans(gctf_ctx* c, int i0, int m, const SearchPlan& sp)
{
	if (sp.ring.n > c->kstride) return fail(c, GCTF_ERR_STATE, "ring larger than the reserved stride");
	return fill_plans_at(c, c->plans.as<ImgPlan>() + i0, m, sp);
}

// Radial average + CFindCtf1D::Do1D / Refine1D for m images: spectra
// [m][T][T/2+1], radial averages [m][T/2+1], plans and states [m]; part
// [m][ldp] holds the workgroup maxima of a brute-force stage
// (ldp >= s1d_part_ld(c)).
int launch_1d_at(gctf_ctx* c, const TileTables& tt, const float* spect, float* radial, const ImgPlan* plans,
	SearchState* st, int m, int mode, float init_df, float range, float2* part, int ldp, int groups)
{
	hipStream_t s = c->stream;
	const int T = tt.T, cx = T / 2 + 1;
	PROF(c, "k_radial", k_radial<<<dim3((cx + 63) / 64, m), 64, 0, s>>>(spect, cx, T, tt.roff.as<int>(),
		tt.rcnt.as<int>(), tt.ridx.as<int>(), radial, cx));
	Search1DParams p{};
	p.cmpx = cx;
	p.mode = mode; p.init_df = init_df; p.range = range;
	p.scan_lo = c->scan_lo; p.scan_hi = c->scan_hi; p.scan_step = c->scan_step;
	PROF(c, "k_s1d_begin", k_s1d_begin<<<(m + 63) / 64, 64, 0, s>>>(p, plans, st, m));
	// Do1D: the scan, then mRefineDefocus around its best; Refine1D: the latter only
	for (int round = 0; round < (mode == 0 ? 2 : 1); ++round) {
		PROF(c, "k_s1d_eval", k_s1d_eval<<<dim3(groups, m), 256, 0, s>>>(cx, plans, radial, tt.w1d.as<double>(), st,
			ldp, part));
		PROF(c, "k_s1d_pick", k_s1d_pick<<<m, 256, 0, s>>>(st, plans, ldp, part));
	}
	HIPCHK(c, hipGetLastError());
	return GCTF_OK;
}

// Radial average + CFindCtf1D::Do1D / Refine1D for images [i0, i0 + m);
// phase_search: the states may hold a phase range (37 phase steps).
int launch_1d(gctf_ctx* c, const TileTables& tt, int i0, int m, int mode, float init_df, float range,
	bool phase_search = true)
{
	const int cx = tt.T / 2 + 1;
	const int ldp = s1d_part_ld(c);
	if ((size_t)(i0 + m) * ldp * sizeof(float2) > c->s1part.bytes)
		return fail(c, GCTF_ERR_STATE, "1D search buffer smaller than the configured scan");
	return launch_1d_at(c, tt, c->spect.as<float>() + (size_t)i0 * tt.T * cx, c->radial.as<float>() + (size_t)i0 * cx,
		c->plans.as<ImgPlan>() + i0, c->states.as<SearchState>() + i0, m, mode, init_df, range,
		c->s1part.as<float2>() + (size_t)i0 * ldp, ldp, s1d_groups(c, phase_search ? 37 : 1));
}

// Device pointers of the 2D search for images [i0, i0 + m).
struct Search2DView {
	const float* spect;
	const ImgPlan* plans;
	float* ring_spec;
	double* ring_ss;
	SearchState* st;
	double2* part;
	RefineArgs* refine;
	float* shell_cc;
	float* ctfmap;   // final-parameter CTF per half-spectrum pixel (the low-pass scratch)
	double* bins;    // spectral-CC sums per integer radius bin
	const int *boff, *bcnt, *bidx;
	int nbins;
	int T, cx, kstride;
};

Search2DView view_2d(gctf_ctx* c, int T, int i0)
{
	Search2DView v;
	v.T = T;
	v.cx = T / 2 + 1;
	v.kstride = c->kstride;
	v.spect = c->spect.as<float>() + (size_t)i0 * T * v.cx;
	v.plans = c->plans.as<ImgPlan>() + i0;
	v.ring_spec = c->ring_spec.as<float>() + (size_t)i0 * c->kstride;
	v.ring_ss = c->ring_ss.as<double>() + i0;
	v.st = c->states.as<SearchState>() + i0;
	v.part = c->part2d.as<double2>() + (size_t)i0 * kMaxCands * kPixChunks;
	v.refine = c->refine.as<RefineArgs>() + i0;
	v.shell_cc = c->shell_cc.as<float>() + (size_t)i0 * v.cx;
	v.ctfmap = c->q.as<float>() + (size_t)i0 * T * v.cx;
	const TileTables& tt = *c->tiles.at(T);
	v.nbins = tt.nbins;
	v.boff = tt.boff.as<int>();
	v.bcnt = tt.bcnt.as<int>();
	v.bidx = tt.bidx.as<int>();
	v.bins = c->shell_bins.as<double>() + (size_t)i0 * 6 * tt.nbins;
	return v;
}

// A 2D search program: `begin`, INIT, the astigmatism grid, `iters` - 1
// rounds of (ASTMAG, ASTANG, DFMEAN, PHASE), the optional fine defocus
// stage, then mCalcCtfRes.  The plain program is CFindDefocus2D::mIterate
// (CFindDefocus2D.cpp:153-176) with the reference's counts.
int run_2d_program(gctf_ctx* c, const Search2DView& v, int m, int begin, const SearchProgram& pg,
	float phase_range)
{
	hipStream_t s = c->stream;
	PROF(c, "k_ring_gather", k_ring_gather<<<m, 1024, 0, s>>>(v.spect, v.T * v.cx, v.plans, v.kstride,
		v.ring_spec, v.ring_ss));
	// search blocking: candidates per workgroup (each candidate's sums do not
	// depend on its neighbours, so the choice leaves the results unchanged)
	auto eval = [&](int maxc) {
		const int cpb = c->cpb;
		dim3 grid((maxc + cpb - 1) / cpb, kPixChunks, m);
		if (cpb == 2)
			PROF(c, "k_cc2d_eval", k_cc2d_eval<2><<<grid, 256, 0, s>>>(v.plans, v.ring_spec, v.kstride, v.st, v.part));
		else if (cpb == 4)
			PROF(c, "k_cc2d_eval", k_cc2d_eval<4><<<grid, 256, 0, s>>>(v.plans, v.ring_spec, v.kstride, v.st, v.part));
		else if (cpb == 16)
			PROF(c, "k_cc2d_eval", k_cc2d_eval<16><<<grid, 256, 0, s>>>(v.plans, v.ring_spec, v.kstride, v.st, v.part));
		else
			PROF(c, "k_cc2d_eval", k_cc2d_eval<8><<<grid, 256, 0, s>>>(v.plans, v.ring_spec, v.kstride, v.st, v.part));
	};
	auto step = [&](int next, int next2 = -1) {
		PROF(c, "k_s2d_step", k_s2d_step<<<m, 256, 0, s>>>(v.st, v.part, v.ring_ss, next, next2, v.refine, pg));
	};
	int maxph = 61;
	if (phase_range > 61.0f) maxph = std::min(kMaxCands, (int)phase_range + 3);
	const int last = pg.fine_n >= 2 ? NX_FINE : NX_FINISH;
	step(begin);
	eval(1);
	step(NX_ASTIG);
	eval(std::min(kMaxCands, pg.n_ast * pg.n_ang));
	step(NX_ASTMAG);
	for (int i = 1; i < pg.iters; ++i) {
		// line stages hold at most n + 1 candidates (odd counts, prep_line / prep_dfmean)
		eval(pg.n_mag + 1);
		step(NX_ASTANG);
		eval(pg.n_angl + 1);
		step(NX_DFMEAN);
		eval(pg.n_df + 1);
		const int after = i < pg.iters - 1 ? NX_ASTMAG : last;
		if (phase_range > 0.0f) {
			step(NX_PHASE);
			eval(maxph);
			step(after);
		} else {
			step(NX_PHASE, after);   // the phase stage prepares nothing
		}
	}
	if (last == NX_FINE) {
		eval(pg.fine_n + 1);
		step(NX_FINISH);
	}
	PROF(c, "k_ctf_map", k_ctf_map<<<dim3(v.T, m), 256, 0, s>>>(v.plans, v.cx, v.T, v.st, v.ctfmap));
	PROF(c, "k_shell_bins", k_shell_bins<<<dim3(v.nbins, m), 64, 0, s>>>(v.spect, v.ctfmap, v.T * v.cx, v.boff,
		v.bcnt, v.bidx, v.nbins, v.st, v.bins));
	PROF(c, "k_shell_cc", k_shell_cc<<<dim3((v.cx + 63) / 64, m), 64, 0, s>>>(v.bins, v.nbins, v.cx, 10, v.st,
		v.shell_cc));
	PROF(c, "k_ctf_res", k_ctf_res<<<(m + 63) / 64, 64, 0, s>>>(v.shell_cc, v.cx, v.T, v.plans, v.st, m));
	HIPCHK(c, hipGetLastError());
	return GCTF_OK;
}

// CFindDefocus2D::DoIt / Refine (plain program) for images [i0, i0 + m).
int launch_2d(gctf_ctx* c, int T, int i0, int m, int begin, float phase_range)
{
	return run_2d_program(c, view_2d(c, T, i0), m, begin, plain_program(), phase_range);
}

// ---- learned path --------------------------------------------------------
// Network outputs -> proposal (mean defocus A, astigmatism A, azimuth deg,
// confidence) and the defocus window of the narrowed search.  A micrograph
// whose confidence is below conf_thr goes to the plain search whatever the
// narrowed search finds, so it sits the narrowed search out (skip).
__global__ void k_proposal(const float* __restrict__ o, int n, float window, float conf_thr,
	float* __restrict__ prop, RefineArgs* __restrict__ refine, SearchState* __restrict__ st)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const float* oi = o + i * kNetOut;
	const float df = 1.0e4f * expf(fminf(fmaxf(oi[0], -3.0f), 3.0f));
	float az = oi[2] * 57.29577951308232f;
	az = az - 180.0f * floorf((az + 90.0f) / 180.0f);   // [-90, 90)
	const float conf = 1.0f / (1.0f + expf(-oi[3]));
	prop[i * 4 + 0] = df;
	prop[i * 4 + 1] = 1.0e3f * oi[1];
	prop[i * 4 + 2] = az;
	prop[i * 4 + 3] = conf;
	RefineArgs r;
	r.dfmean[0] = df; r.dfmean[1] = 2.0f * window;
	r.ast[0] = 0.0f; r.ast[1] = 0.0f;
	r.ang[0] = 0.0f; r.ang[1] = 0.0f;
	r.phase[0] = 0.0f; r.phase[1] = 0.0f;
	refine[i] = r;
	st[i].conf = conf;
	st[i].skip = conf < conf_thr ? 1 : 0;
}

// Hand the micrograph to the plain search when the proposal is not trusted
// (confidence below the threshold), the narrowed search failed (score <= 0,
// the reference's failure signal), or its defocus ended within one fine step
// of the window edge (the optimum may lie outside).
__global__ void k_learned_check(SearchState* __restrict__ st, int n, float conf_thr)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	SearchState& s = st[i];
	const float edge = fmaxf(s.pg.fine_step, 1.0f);
	const bool at_lo = s.dfrange[0] > 3000.0f && s.dfmean - s.dfrange[0] < edge;
	const bool at_hi = s.dfrange[1] - s.dfmean < edge;
	s.fallback = (s.conf < conf_thr || !(s.score > 0.0f) || at_lo || at_hi) ? 1 : 0;
}

// Before the fallback pass: images the learned path settled sit out; the
// others restart from CFindCtfBase::SetPhase's values.
__global__ void k_fallback_prepare(SearchState* __restrict__ st, int n, float p0, float range)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	SearchState& s = st[i];
	if (!s.fallback) { s.skip = 1; return; }
	s.skip = 0;
	s.ext_phase = p0;
	s.phase_range = range;
}

__global__ void k_clear_skip(SearchState* __restrict__ st, int n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) st[i].skip = 0;
}

__global__ void k_state_init(SearchState* st, int n, float p0, float range)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	SearchState s;
	memset(&s, 0, sizeof s);
	s.pg = plain_program();
	s.ext_phase = p0;          // CFindCtfBase::SetPhase
	s.phase_range = range;
	st[i] = s;
}

// Search results in the ABI's layout, written on the device.
__global__ void k_results_out(const SearchState* __restrict__ st, int n, gctf_result* __restrict__ out)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const SearchState& s = st[i];
	gctf_result r;
	r.df_min = s.df_min;
	r.df_max = s.df_max;
	r.azimuth_deg = s.ast_ang;
	r.ext_phase_deg = s.ext_phase;
	r.score = s.score;
	r.ctf_res = s.ctf_res;
	out[i] = r;
}

int read_results(gctf_ctx* c, int n, gctf_result* out)
{
	HIPCHK(c, hipMemcpyAsync(c->h_pinned, c->states.p, sizeof(SearchState) * n, hipMemcpyDeviceToHost,
		c->stream));
	HIPCHK(c, hipStreamSynchronize(c->stream));
	for (int i = 0; i < n; ++i) {
		const SearchState& s = c->h_pinned[i];
		out[i].df_min = s.df_min;
		out[i].df_max = s.df_max;
		out[i].azimuth_deg = s.ast_ang;
		out[i].ext_phase_deg = s.ext_phase;
		out[i].score = s.score;
		out[i].ctf_res = s.ctf_res;
	}
	return GCTF_OK;
}

// ---- diagnostic full spectrum (CFindCtfBase::GenFullSpectrum) -------------
__global__ __launch_bounds__(256) void k_moments1(const float* __restrict__ sp, int n, double* out)
{
	__shared__ double r1[256], r2[256];
	double a = 0.0, b = 0.0;
	for (int i = threadIdx.x; i < n; i += 256) { double v = sp[i]; a += v; b += v * v; }
	r1[threadIdx.x] = a; r2[threadIdx.x] = b;
	__syncthreads();
	for (int o = 128; o > 0; o >>= 1) {
		if (threadIdx.x < o) { r1[threadIdx.x] += r1[threadIdx.x + o]; r2[threadIdx.x] += r2[threadIdx.x + o]; }
		__syncthreads();
	}
	if (threadIdx.x == 0) { out[0] = r1[0] / n; out[1] = r2[0] / n; }
}

// CSpectrumImage::DoIt (CSpectrumImage.cpp:21-71): unpadded unfold
// (GCalcSpectrum.cu:47-69, row cmpy taken mod cmpy) and |CTF| painted on the
// left half (GCalcCTF2D.cu:49-78).  Grid (T), 256 threads.
__global__ __launch_bounds__(256) void k_full_spectrum(const float* __restrict__ half, int cmpx, int cmpy,
	CtfConsts cc, float dfmin_pix, float dfmax_pix, float az, float add, float fmin, float fmax,
	const double* __restrict__ mom, float* __restrict__ out)
{
	const int hx = cmpx - 1, nxf = 2 * hx, y = blockIdx.x;
	const double mean = mom[0], var = mom[1] - mom[0] * mom[0];
	const float sd = var < 0 ? 0.0f : (float)sqrt(var);
	const float gain = sd * 1.5f;
	const float dmean = 0.5f * (dfmin_pix + dfmax_pix), dsig = 0.5f * (dfmax_pix - dfmin_pix);
	const float w2 = cc.wl * cc.wl;
	for (int xd = threadIdx.x; xd < nxf; xd += 256) {
		int xs = xd - hx, ys = y;
		if (xs < 0) { xs = -xs; ys = (cmpy - y) % cmpy; }
		float v = half[ys * cmpx + xs];
		if (xd < hx) {
			float fy = ((float)y - (float)cmpy * 0.5f) / (float)cmpy;
			float fx = ((float)xd - (float)hx) * 0.5f / (float)hx;
			fx = sqrtf(fx * fx + fy * fy);
			if (!(fx < fmin || fx > fmax)) {
				int ix = hx - xd, iy = (cmpy - y) % cmpy;
				float fX = (float)ix * 0.5f / (float)(cmpx - 1);
				float fY = (float)(iy - cmpy / 2) / (float)cmpy;
				float s2 = fX * fX + fY * fY;
				float th = atanf(fY / (fX + (float)1e-30));
				float df = dmean + dsig * cosf(2.0f * (th - az));
				float ctf = -sinf(add + kPi2D * cc.wl * s2 * (df - 0.5f * s2 * w2 * cc.cs));
				v = (float)(((double)fabsf(ctf) - 0.5) * gain + mean);
			}
		}
		out[(long long)y * nxf + xd] = v;
	}
}

int need(gctf_ctx* c, bool cond, const char* what)
{
	if (!cond) return fail(c, GCTF_ERR_STATE, what);
	return GCTF_OK;
}

int current_plan(gctf_ctx* c, SearchPlan** sp, TileTables** tt)
{
	CHK(need(c, c->setup1, "Setup1 has not been called"));
	CHK(get_search_plan(c, c->s_kv, c->s_cs, c->s_ac, c->s_pix, c->prm.tile_size, sp));
	CHK(get_tile_tables(c, c->prm.tile_size, tt));
	CHK(ensure_image_buffers(c, 1, c->prm.tile_size, (*sp)->ring.n));
	c->kstride = (*sp)->ring.n;
	CHK(fill_plans(c, 0, 1, **sp));
	return GCTF_OK;
}

int single_result(gctf_ctx* c, gctf_result* out)
{
	gctf_result tmp;
	CHK(read_results(c, 1, &tmp));
	if (out) *out = tmp;
	return GCTF_OK;
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

int gctf_create(const gctf_params* params, int device, gctf_ctx** out)
{
	if (!out) return GCTF_ERR_ARG;
	*out = nullptr;
	gctf_params p;
	if (params) p = *params;
	else { p.kv = 300.f; p.cs_mm = 2.7f; p.amp_contrast = 0.07f; p.ext_phase_deg = 0.f;
		p.ext_phase_range_deg = 0.f; p.tile_size = 512; p.log_spect = 0; }
	if (!(p.kv > 0) || !(p.amp_contrast >= 0 && p.amp_contrast < 1)) return GCTF_ERR_ARG;
	if (p.tile_size != 256 && p.tile_size != 512 && p.tile_size != 1024) return GCTF_ERR_UNSUPPORTED;
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return GCTF_ERR_HIP;
	if (hipSetDevice(device) != hipSuccess) return GCTF_ERR_HIP;
	auto c = std::make_unique<gctf_ctx>();
	c->device = device;
	c->prm = p;
	if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) return GCTF_ERR_HIP;
	c->stream = c->own_stream;
	// this device's static scratch (gctf_do2d_batch_static); a failed lookup
	// disables that entry point only
	const hipError_t se = static_scratch_address(device, &c->static_scratch);
	if (se != hipSuccess) {
		c->static_scratch = nullptr;
		c->static_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(se);
		(void)hipGetLastError();
	}
	*out = c.release();
	return GCTF_OK;
}

void gctf_destroy(gctf_ctx* ctx)
{
	if (!ctx) return;
	(void)hipSetDevice(ctx->device);
	(void)hipStreamSynchronize(ctx->stream);
	delete ctx;
}

const char* gctf_error_string(const gctf_ctx* ctx)
{
	if (!ctx) return "null context";
	return ctx->err.c_str();
}

int gctf_set_stream(gctf_ctx* ctx, void* hip_stream)
{
	if (!ctx) return GCTF_ERR_ARG;
	ctx->stream = hip_stream ? (hipStream_t)hip_stream : ctx->own_stream;
	for (auto& kv : ctx->graphs)
		if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
	ctx->graphs.clear();
	return GCTF_OK;
}

int gctf_synchronize(gctf_ctx* ctx)
{
	if (!ctx) return GCTF_ERR_ARG;
	HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
	return GCTF_OK;
}

int gctf_host_alloc(size_t bytes, void** out)
{
	if (!out) return GCTF_ERR_ARG;
	*out = nullptr;
	if (bytes == 0) return GCTF_ERR_ARG;
	void* p = nullptr;
	if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess || !p) {
		(void)hipGetLastError();
		return GCTF_ERR_NOMEM;
	}
	*out = p;
	return GCTF_OK;
}

void gctf_host_free(void* p)
{
	if (p) (void)hipHostFree(p);
}

int gctf_tune(gctf_ctx* ctx, int chunk, int tile_splits, int use_graph, int search_block, int list_queue)
{
	if (!ctx) return GCTF_ERR_ARG;
	if (search_block != 0 && search_block != 2 && search_block != 4 && search_block != 8 && search_block != 16)
		return fail(ctx, GCTF_ERR_ARG, "search blocking: 2, 4, 8 or 16 candidates per workgroup");
	if (chunk > 0) ctx->chunk = chunk;
	if (tile_splits > 0) { ctx->tile_splits = tile_splits; ctx->gen++; }
	if (use_graph != 0) ctx->use_graph = use_graph > 0 ? 1 : 0;
	if (search_block != 0) ctx->cpb = search_block;
	if (list_queue != 0) ctx->list_queue = list_queue > 0 ? 1 : 0;
	return GCTF_OK;
}

int gctf_profile(gctf_ctx* ctx, int enable)
{
	if (!ctx) return GCTF_ERR_ARG;
	HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
	for (auto& r : ctx->prec) { ctx->evpool.push_back(r.a); ctx->evpool.push_back(r.b); }
	ctx->prec.clear();
	ctx->ptotal.clear();
	ctx->prof = enable != 0;
	return GCTF_OK;
}

int gctf_profile_report(gctf_ctx* ctx, int max, char* names, double* total_ms, int* count, int* nout)
{
	if (!ctx || !nout) return GCTF_ERR_ARG;
	HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
	for (auto& r : ctx->prec) {
		float ms = 0.f;
		HIPCHK(ctx, hipEventElapsedTime(&ms, r.a, r.b));
		auto& t = ctx->ptotal[r.name];
		t.first += ms;
		t.second += 1;
		ctx->evpool.push_back(r.a);
		ctx->evpool.push_back(r.b);
	}
	ctx->prec.clear();
	int i = 0;
	for (auto& kv : ctx->ptotal) {
		if (i >= max) break;
		if (names) {
			strncpy(names + 64 * i, kv.first.c_str(), 63);
			names[64 * i + 63] = 0;
		}
		if (total_ms) total_ms[i] = kv.second.first;
		if (count) count[i] = kv.second.second;
		++i;
	}
	*nout = i;
	return GCTF_OK;
}

int gctf_rescale_setup(gctf_ctx* ctx, int nx, int ny, float raw_pix, int* new_size, float* new_pix)
{
	if (!ctx) return GCTF_ERR_ARG;
	HIPCHK(ctx, hipSetDevice(ctx->device));
	ctx->rp_ready = false;
	ctx->rp_done = false;
	CHK(make_rescale_plan(ctx, nx, ny, raw_pix, ctx->rp));
	CHK(ensure_buf(ctx, ctx->pad1, sizeof(float) * (size_t)ctx->rp.pitch * ctx->rp.nny));
	if (ctx->rp.active) CHK(ensure_rescale_buffers(ctx, ctx->rp, 1));
	ctx->rp_ready = true;
	if (new_size) { new_size[0] = ctx->rp.nnx; new_size[1] = ctx->rp.nny; new_size[2] = ctx->rp.pitch; }
	if (new_pix) *new_pix = ctx->rp.pixn;
	return GCTF_OK;
}

// The device copy gctf_rescale_prefetch made of host image img (bytes long),
// if the last prefetch was of that image: the context's stream waits for it,
// and it becomes the staging buffer.  Null otherwise.  A prefetch serves
// only the next rescale of a host image and is dropped by any other.
static const float* take_prefetch(gctf_ctx* ctx, const float* img, size_t bytes, int* rc)
{
	*rc = GCTF_OK;
	const bool match = ctx->prefetch_src == img && ctx->prefetch_bytes == bytes;
	ctx->prefetch_src = nullptr;
	if (!match) return nullptr;
	const hipError_t e = hipStreamWaitEvent(ctx->stream, ctx->prefetch_done, 0);
	if (e != hipSuccess) {
		*rc = fail(ctx, GCTF_ERR_HIP, std::string("hipStreamWaitEvent: ") + hipGetErrorString(e));
		return nullptr;
	}
	std::swap(ctx->hostin.p, ctx->prefetch.p);
	std::swap(ctx->hostin.bytes, ctx->prefetch.bytes);
	return ctx->hostin.as<float>();
}

int gctf_rescale(gctf_ctx* ctx, const float* img, int img_on_device)
{
	if (!ctx || !img) return fail(ctx, GCTF_ERR_ARG, "null image");
	CHK(need(ctx, ctx->rp_ready, "CRescaleImage::Setup has not been called"));
	HIPCHK(ctx, hipSetDevice(ctx->device));
	const RescalePlan& rp = ctx->rp;
	float* out = ctx->pad1.as<float>();
	const size_t bytes = sizeof(float) * (size_t)rp.nx * rp.ny;
	int rc = GCTF_OK;
	const float* pre = img_on_device ? nullptr : take_prefetch(ctx, img, bytes, &rc);
	if (img_on_device) ctx->prefetch_src = nullptr;
	CHK(rc);
	if (!rp.active) {
		// CRescaleImage.cpp:76-85: rows copied into the padded buffer
		const bool dev = img_on_device || pre;
		HIPCHK(ctx, hipMemcpy2DAsync(out, sizeof(float) * rp.pitch, pre ? pre : img, sizeof(float) * rp.nx,
			sizeof(float) * rp.nx, rp.nny, dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->stream));
	} else {
		const float* dimg = img;
		if (pre) {
			dimg = pre;
		} else if (!img_on_device) {
			// the staging buffer stays with the context: no allocation, and no
			// synchronising free, per image (no graph reads it, so growing it
			// leaves the cached graphs valid)
			const hipError_t e = ctx->hostin.ensure(bytes, nullptr);
			if (e != hipSuccess)
				return fail(ctx, GCTF_ERR_NOMEM, std::string("device allocation of ") + std::to_string(bytes)
					+ " bytes failed: " + hipGetErrorString(e));
			HIPCHK(ctx, hipMemcpyAsync(ctx->hostin.p, img, bytes, hipMemcpyHostToDevice, ctx->stream));
			dimg = ctx->hostin.as<float>();
		}
		CHK(launch_rescale(ctx, rp, dimg, GCTF_DTYPE_F32, (long long)rp.nx * rp.ny, 1, out,
			(long long)rp.pitch * rp.nny));
	}
	HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
	ctx->rp_done = true;
	return GCTF_OK;
}

int gctf_rescale_prefetch(gctf_ctx* ctx, const float* img, int nx, int ny)
{
	if (!ctx || !img || nx < 1 || ny < 1) return fail(ctx, GCTF_ERR_ARG, "null image or empty size");
	HIPCHK(ctx, hipSetDevice(ctx->device));
	if (!ctx->copy_stream) HIPCHK(ctx, hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
	if (!ctx->prefetch_done) HIPCHK(ctx, hipEventCreateWithFlags(&ctx->prefetch_done, hipEventDisableTiming));
	if (!ctx->main_done) HIPCHK(ctx, hipEventCreateWithFlags(&ctx->main_done, hipEventDisableTiming));
	const size_t bytes = sizeof(float) * (size_t)nx * ny;
	ctx->prefetch_src = nullptr;
	// the buffer may last have been read by work on the context's stream
	HIPCHK(ctx, hipEventRecord(ctx->main_done, ctx->stream));
	HIPCHK(ctx, hipStreamWaitEvent(ctx->copy_stream, ctx->main_done, 0));
	const hipError_t e = ctx->prefetch.ensure(bytes, nullptr);
	if (e != hipSuccess)
		return fail(ctx, GCTF_ERR_NOMEM, std::string("device allocation of ") + std::to_string(bytes)
			+ " bytes failed: " + hipGetErrorString(e));
	HIPCHK(ctx, hipMemcpyAsync(ctx->prefetch.p, img, bytes, hipMemcpyHostToDevice, ctx->copy_stream));
	HIPCHK(ctx, hipEventRecord(ctx->prefetch_done, ctx->copy_stream));
	ctx->prefetch_src = img;
	ctx->prefetch_bytes = bytes;
	return GCTF_OK;
}

int gctf_scaled_image(gctf_ctx* ctx, float** dev_pad_img)
{
	if (!ctx || !dev_pad_img) return GCTF_ERR_ARG;
	CHK(need(ctx, ctx->rp_ready, "CRescaleImage::Setup has not been called"));
	*dev_pad_img = ctx->pad1.as<float>();
	return GCTF_OK;
}

int gctf_setup1(gctf_ctx* ctx, float kv, float cs_mm, float amp_contrast, float pix, float ast_tol,
	float ext_phase_rad)
{
	(void)ast_tol;        // CCTFTheory's AstTol is not used by the search
	(void)ext_phase_rad;  // kept in CCTFTheory; the search phase comes from SetPhase
	if (!ctx) return GCTF_ERR_ARG;
	if (!(kv > 0) || !(pix > 0) || !(amp_contrast >= 0 && amp_contrast < 1))
		return fail(ctx, GCTF_ERR_ARG, "bad CTF parameters");
	HIPCHK(ctx, hipSetDevice(ctx->device));
	ctx->s_kv = kv; ctx->s_cs = cs_mm; ctx->s_ac = amp_contrast; ctx->s_pix = pix;
	ctx->setup1 = true;
	ctx->have_raw = ctx->have_spect = false;
	SearchPlan* sp;
	TileTables* tt;
	CHK(current_plan(ctx, &sp, &tt));
	// a fresh CFindDefocus2D (CFindCtf2D::Setup1): astigmatism, angle and its
	// extra phase start at 0; CFindCtfBase's fields persist.
	HIPCHK(ctx, hipMemcpyAsync(ctx->h_pinned, ctx->states.p, sizeof(SearchState), hipMemcpyDeviceToHost, ctx->stream));
	HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
	SearchState s = ctx->h_pinned[0];
	s.ast = 0.f; s.ang = 0.f; s.phase = 0.f; s.done = 0; s.st_kind = ST_NONE; s.st_n = 0;
	ctx->h_pinned[0] = s;
	HIPCHK(ctx, hipMemcpyAsync(ctx->states.p, ctx->h_pinned, sizeof(SearchState), hipMemcpyHostToDevice, ctx->stream));
	HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
	return GCTF_OK;
}

int gctf_setup2(gctf_ctx* ctx, int nx, int ny)
{
	if (!ctx) return GCTF_ERR_ARG;
	CHK(need(ctx, ctx->setup1, "Setup1 has not been called"));
	const int T = ctx->prm.tile_size;
	if (nx < T || ny < T) return fail(ctx, GCTF_ERR_ARG, "image smaller than one tile");
	ctx->s_nx = nx; ctx->s_ny = ny;
	ctx->setup2 = true;
	return GCTF_OK;
}

int gctf_set_search1d(gctf_ctx* ctx, float df_lo_A, float df_hi_A, float step_A)
{
	if (!ctx) return GCTF_ERR_ARG;
	const bool dflt = df_lo_A == 0.0f && df_hi_A == 0.0f && step_A == 0.0f;
	if (!dflt && !(df_lo_A > 0.0f && df_hi_A > df_lo_A && step_A > 0.0f))
		return fail(ctx, GCTF_ERR_ARG, "1D scan: need 0 < df_lo < df_hi and step > 0 (or all 0)");
	if (!dflt && (df_hi_A - df_lo_A) / step_A > 1e6f)
		return fail(ctx, GCTF_ERR_ARG, "1D scan: more than 1e6 points");
	ctx->scan_lo = df_lo_A; ctx->scan_hi = df_hi_A; ctx->scan_step = step_A;
	ctx->gen++;   // recorded graphs hold the old scan
	return GCTF_OK;
}

int gctf_set_phase(gctf_ctx* ctx, float p0_deg, float range_deg)
{
	if (!ctx) return GCTF_ERR_ARG;
	CHK(need(ctx, ctx->setup1, "Setup1 has not been called"));
	HIPCHK(ctx, hipSetDevice(ctx->device));
	HIPCHK(ctx, hipMemcpyAsync(ctx->h_pinned, ctx->states.p, sizeof(SearchState), hipMemcpyDeviceToHost, ctx->stream));
	HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
	ctx->h_pinned[0].ext_phase = p0_deg;
	ctx->h_pinned[0].phase_range = range_deg;
	HIPCHK(ctx, hipMemcpyAsync(ctx->states.p, ctx->h_pinned, sizeof(SearchState), hipMemcpyHostToDevice, ctx->stream));
	HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
	return GCTF_OK;
}

int gctf_gen_half_spectrum(gctf_ctx* ctx, const float* dev_pad_img)
{
	if (!ctx || !dev_pad_img) return fail(ctx, GCTF_ERR_ARG, "null image");
	CHK(need(ctx, ctx->setup2, "Setup2 has not been called"));
	HIPCHK(ctx, hipSetDevice(ctx->device));
	SearchPlan* sp;
	TileTables* tt;
	CHK(current_plan(ctx, &sp, &tt));
	const int T = ctx->prm.tile_size;
	TileGrid g = make_grid(ctx->s_nx, ctx->s_ny, T);
	if (g.ntx < 1 || g.nty < 1) return fail(ctx, GCTF_ERR_ARG, "no tile fits the image");
	CHK(ensure_spectrum_buffers(ctx, g, T, 1));
	const int pitch = (ctx->s_nx / 2 + 1) * 2;   // CGenAvgSpectrum.cpp:130
	CHK(launch_spectrum<float>(ctx, g, *tt, dev_pad_img, pitch, 0, 1, 0));
	CHK(launch_post(ctx, *tt, 1));
	HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
	ctx->have_raw = ctx->have_spect = true;
	return GCTF_OK;
}

int gctf_do1d(gctf_ctx* ctx, gctf_result* out)
{
	if (!ctx) return GCTF_ERR_ARG;
	CHK(need(ctx, ctx->have_spect, "no spectrum: call GenHalfSpectrum or SetHalfSpect"));
	HIPCHK(ctx, hipSetDevice(ctx->device));
	SearchPlan* sp;
	TileTables* tt;
	CHK(current_plan(ctx, &sp, &tt));
	CHK(launch_1d(ctx, *tt, 0, 1, 0, 0.f, 0.f));
	return single_result(ctx, out);
}

int gctf_refine1d(gctf_ctx* ctx, float init_df, float df_range, gctf_result* out)
{
	if (!ctx) return GCTF_ERR_ARG;
	CHK(need(ctx, ctx->have_spect, "no spectrum: call GenHalfSpectrum or SetHalfSpect"));
	HIPCHK(ctx, hipSetDevice(ctx->device));
	SearchPlan* sp;
	TileTables* tt;
	CHK(current_plan(ctx, &sp, &tt));
	CHK(launch_1d(ctx, *tt, 0, 1, 1, init_df, df_range));
	return single_result(ctx, out);
}

int gctf_do2d(gctf_ctx* ctx, gctf_result* out)
{
	if (!ctx) return GCTF_ERR_ARG;
	CHK(need(ctx, ctx->have_spect, "no spectrum: call GenHalfSpectrum or SetHalfSpect"));
	HIPCHK(ctx, hipSetDevice(ctx->device));
	SearchPlan* sp;
	TileTables* tt;
	CHK(current_plan(ctx, &sp, &tt));
	// the phase range in use is the one SetPhase stored in the state
	HIPCHK(ctx, hipMemcpyAsync(ctx->h_pinned, ctx->states.p, sizeof(SearchState), hipMemcpyDeviceToHost, ctx->stream));
	HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
	const float prange = ctx->h_pinned[0].phase_range;
	CHK(launch_1d(ctx, *tt, 0, 1, 0, 0.f, 0.f));
	CHK(launch_2d(ctx, ctx->prm.tile_size, 0, 1, NX_BEGIN_DO2D, prange));
	return single_result(ctx, out);
}

int gctf_do2d_batch_static(gctf_ctx* ctx, const float* d_spect, int n, gctf_result* d_out)
{
	if (!ctx) return GCTF_ERR_ARG;
	if (!d_spect || !d_out || n < 1) return fail(ctx, GCTF_ERR_ARG, "null spectra/output or n < 1");
	if (n > kStaticMaxImages)
		return fail(ctx, GCTF_ERR_ARG, "gctf_do2d_batch_static takes at most " + std::to_string(kStaticMaxImages)
			+ " spectra per call (the size of its static scratch), got " + std::to_string(n));
	if (ctx->prm.tile_size != kStaticTile)
		return fail(ctx, GCTF_ERR_UNSUPPORTED, "gctf_do2d_batch_static needs tile size 512");
	CHK(need(ctx, ctx->setup1, "Setup1 has not been called"));
	if (ctx->prof)
		return fail(ctx, GCTF_ERR_STATE, "kernel profiling is on (it creates events); turn it off for "
			"gctf_do2d_batch_static");
	if (!ctx->static_scratch)
		return fail(ctx, GCTF_ERR_HIP, ctx->static_err.empty() ? "no static scratch" : ctx->static_err);
	HIPCHK(ctx, hipSetDevice(ctx->device));
	// Setup1 made the plan of its parameters and the tile tables: looked up
	// here, never created
	SearchPlan* sp = find_search_plan(ctx, ctx->s_kv, ctx->s_cs, ctx->s_ac, ctx->s_pix, kStaticTile);
	auto tit = ctx->tiles.find(kStaticTile);
	if (!sp || tit == ctx->tiles.end()) return fail(ctx, GCTF_ERR_STATE, "no search plan for Setup1's parameters");
	const TileTables& tt = *tit->second;
	if ((sp->ring.n + 15) / 16 * 16 > kStaticMaxRing)
		return fail(ctx, GCTF_ERR_UNSUPPORTED, "correlation ring of " + std::to_string(sp->ring.n)
			+ " pixels above the static scratch's " + std::to_string(kStaticMaxRing));
	if (tt.nbins > kStaticMaxBins) return fail(ctx, GCTF_ERR_UNSUPPORTED, "radius bins above the static scratch");
	if (s1d_part_ld(ctx) > kStaticS1dLd)
		return fail(ctx, GCTF_ERR_UNSUPPORTED, "configured 1D scan longer than the static scratch's 501 points");
	char* base = static_cast<char*>(ctx->static_scratch);
	SearchState* st = reinterpret_cast<SearchState*>(base + kStStates);
	ImgPlan* plans = reinterpret_cast<ImgPlan*>(base + kStPlans);
	float* radial = reinterpret_cast<float*>(base + kStRadial);
	Search2DView v;
	v.T = kStaticTile;
	v.cx = kStaticCx;
	v.kstride = kStaticMaxRing;
	v.spect = d_spect;
	v.plans = plans;
	v.ring_spec = reinterpret_cast<float*>(base + kStRing);
	v.ring_ss = reinterpret_cast<double*>(base + kStRingSs);
	v.st = st;
	v.part = reinterpret_cast<double2*>(base + kStPart);
	v.refine = reinterpret_cast<RefineArgs*>(base + kStRefine);
	v.shell_cc = reinterpret_cast<float*>(base + kStShellCc);
	v.ctfmap = reinterpret_cast<float*>(base + kStCtfMap);
	v.bins = reinterpret_cast<double*>(base + kStBins);
	v.nbins = tt.nbins;
	v.boff = tt.boff.as<int>();
	v.bcnt = tt.bcnt.as<int>();
	v.bidx = tt.bidx.as<int>();
	const float p0 = ctx->prm.ext_phase_deg, prange = std::max(ctx->prm.ext_phase_range_deg, 0.0f);
	// the array keeps the previous call's contents: every state starts fresh
	// (the other buffers are written before they are read)
	k_state_init<<<(n + 63) / 64, 64, 0, ctx->stream>>>(st, n, p0, prange);
	HIPCHK(ctx, hipGetLastError());
	CHK(fill_plans_at(ctx, plans, n, *sp));
	CHK(launch_1d_at(ctx, tt, d_spect, radial, plans, st, n, 0, 0.f, 0.f, reinterpret_cast<float2*>(base + kStS1d),
		kStaticS1dLd, s1d_groups(ctx, prange > 0.0f ? 37 : 1)));
	CHK(run_2d_program(ctx, v, n, NX_BEGIN_DO2D, plain_program(), prange));
	k_results_out<<<(n + 63) / 64, 64, 0, ctx->stream>>>(st, n, d_out);
	HIPCHK(ctx, hipGetLastError());
	return GCTF_OK;
}

int gctf_refine(gctf_ctx* ctx, const float* df_mean, const float* ast_ratio, const float* ast_angle_deg,
	const float* ext_phase_deg, gctf_result* out)
{
	if (!ctx || !df_mean || !ast_ratio || !ast_angle_deg || !ext_phase_deg)
		return fail(ctx, GCTF_ERR_ARG, "null refine range");
	CHK(need(ctx, ctx->have_spect, "no spectrum: call GenHalfSpectrum or SetHalfSpect"));
	HIPCHK(ctx, hipSetDevice(ctx->device));
	SearchPlan* sp;
	TileTables* tt;
	CHK(current_plan(ctx, &sp, &tt));
	RefineArgs r;
	for (int i = 0; i < 2; ++i) {
		r.dfmean[i] = df_mean[i]; r.ast[i] = ast_ratio[i];
		r.ang[i] = ast_angle_deg[i]; r.phase[i] = ext_phase_deg[i];
	}
	HIPCHK(ctx, hipMemcpyAsync(ctx->refine.p, &r, sizeof r, hipMemcpyHostToDevice, ctx->stream));
	CHK(launch_2d(ctx, ctx->prm.tile_size, 0, 1, NX_BEGIN_REFINE, ext_phase_deg[1]));
	return single_result(ctx, out);
}

int gctf_spect_size(gctf_ctx* ctx, int* size, int half)
{
	if (!ctx || !size) return GCTF_ERR_ARG;
	const int T = ctx->prm.tile_size;
	size[0] = half ? T / 2 + 1 : T;   // CFindCtfBase.cpp:101-107
	size[1] = T;
	return GCTF_OK;
}

int gctf_get_half_spect(gctf_ctx* ctx, int raw, float* host_out)
{
	if (!ctx || !host_out) return fail(ctx, GCTF_ERR_ARG, "null output");
	CHK(need(ctx, raw ? ctx->have_raw : ctx->have_spect, "no spectrum"));
	HIPCHK(ctx, hipSetDevice(ctx->device));
	const int T = ctx->prm.tile_size;
	const size_t n = (size_t)(T / 2 + 1) * T;
	HIPCHK(ctx, hipMemcpyAsync(host_out, raw ? ctx->raw.p : ctx->spect.p, sizeof(float) * n,
		hipMemcpyDeviceToHost, ctx->stream));
	HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
	return GCTF_OK;
}

int gctf_set_half_spect(gctf_ctx* ctx, const float* host_spect)
{
	if (!ctx || !host_spect) return fail(ctx, GCTF_ERR_ARG, "null spectrum");
	CHK(need(ctx, ctx->setup1, "Setup1 has not been called"));
	HIPCHK(ctx, hipSetDevice(ctx->device));
	SearchPlan* sp;
	TileTables* tt;
	CHK(current_plan(ctx, &sp, &tt));
	const int T = ctx->prm.tile_size;
	const size_t n = (size_t)(T / 2 + 1) * T;
	HIPCHK(ctx, hipMemcpyAsync(ctx->spect.p, host_spect, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
	HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
	ctx->have_spect = true;
	return GCTF_OK;
}

int gctf_full_spectrum(gctf_ctx* ctx, float* host_out)
{
	if (!ctx || !host_out) return fail(ctx, GCTF_ERR_ARG, "null output");
	CHK(need(ctx, ctx->have_spect, "no spectrum"));
	HIPCHK(ctx, hipSetDevice(ctx->device));
	SearchPlan* sp;
	TileTables* tt;
	CHK(current_plan(ctx, &sp, &tt));
	const int T = ctx->prm.tile_size, cx = T / 2 + 1;
	CHK(ensure_buf(ctx, ctx->fullsp, sizeof(float) * (size_t)T * T));
	CHK(ensure_buf(ctx, ctx->moments1, sizeof(double) * 2));
	HIPCHK(ctx, hipMemcpyAsync(ctx->h_pinned, ctx->states.p, sizeof(SearchState), hipMemcpyDeviceToHost, ctx->stream));
	HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
	const SearchState s = ctx->h_pinned[0];
	// CFindCtfBase::GenFullSpectrum (CFindCtfBase.cpp:118-133) through
	// CCTFTheory::SetDefocus / SetExtPhase (CCTFTheory.cpp:93-118)
	const float pi_f = (float)(4.0 * atan(1.0));
	const float az = s.ast_ang * 0.017453f;
	const float dmin = s.df_min / sp->pix, dmax = s.df_max / sp->pix;
	const float ext = fmodf(s.ext_phase * 0.01745329f, pi_f);
	const float add = sp->cc.amp_phase + ext;
	const float fmin = sp->pix / sp->res[0], fmax = sp->pix / sp->res[1];
	k_moments1<<<1, 256, 0, ctx->stream>>>(ctx->spect.as<float>(), cx * T, ctx->moments1.as<double>());
	k_full_spectrum<<<T, 256, 0, ctx->stream>>>(ctx->spect.as<float>(), cx, T, sp->cc, dmin, dmax, az, add,
		fmin, fmax, ctx->moments1.as<double>(), ctx->fullsp.as<float>());
	HIPCHK(ctx, hipGetLastError());
	HIPCHK(ctx, hipMemcpyAsync(host_out, ctx->fullsp.p, sizeof(float) * (size_t)T * T, hipMemcpyDeviceToHost,
		ctx->stream));
	HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
	return GCTF_OK;
}

int gctf_tile_grid(int nx, int ny, int tile, int* out6)
{
	if (!out6 || nx < 1 || ny < 1 || tile < 2) return GCTF_ERR_ARG;
	const TileGrid g = make_grid(nx, ny, tile);
	out6[0] = g.ntx; out6[1] = g.nty; out6[2] = g.offx; out6[3] = g.offy;
	out6[4] = tile - g.step; out6[5] = g.step;
	return GCTF_OK;
}

int gctf_ring_pixels(gctf_ctx* ctx, float pix, float res_lo, float res_hi, int* idx, int max, int* n)
{
	if (!ctx || !n || !(pix > 0)) return fail(ctx, GCTF_ERR_ARG, "null output or bad pixel size");
	if ((res_lo > 0.f) != (res_hi > 0.f) || (res_lo > 0.f && !(res_lo > res_hi)))
		return fail(ctx, GCTF_ERR_ARG, "resolution bounds: both 0, or res_lo > res_hi > 0");
	HIPCHK(ctx, hipSetDevice(ctx->device));
	SearchPlan* sp;
	CHK(get_search_plan(ctx, ctx->prm.kv, ctx->prm.cs_mm, ctx->prm.amp_contrast, pix, ctx->prm.tile_size, &sp,
		res_lo, res_hi));
	*n = sp->ring.n;
	if (idx && max > 0)
		HIPCHK(ctx, hipMemcpy(idx, sp->idx.p, sizeof(int) * (size_t)std::min(max, sp->ring.n),
			hipMemcpyDeviceToHost));
	return GCTF_OK;
}

int gctf_radial_average(gctf_ctx* ctx, const float* d_spect, int n, float* d_radial, int ld)
{
	if (!ctx || !d_spect || !d_radial || n < 1) return fail(ctx, GCTF_ERR_ARG, "null spectra/output or n < 1");
	const int T = ctx->prm.tile_size, cx = T / 2 + 1;
	if (ld < cx) return fail(ctx, GCTF_ERR_ARG, "ld smaller than tile / 2 + 1");
	HIPCHK(ctx, hipSetDevice(ctx->device));
	TileTables* tt;
	CHK(get_tile_tables(ctx, T, &tt));
	k_radial<<<dim3((cx + 63) / 64, n), 64, 0, ctx->stream>>>(d_spect, cx, T, tt->roff.as<int>(),
		tt->rcnt.as<int>(), tt->ridx.as<int>(), d_radial, ld);
	HIPCHK(ctx, hipGetLastError());
	HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
	return GCTF_OK;
}

}  // extern "C"

// ---- batched path ----------------------------------------------------------
// (the ABI functions below keep C linkage from their declarations in gctf_hip.h)
namespace {

// Search variants of a batch after the shared spectrum stages.
struct BatchJob {
	int mode = 0;                 // 0: Do1D + Do2D, 1: BF16 grid search, 2: learned, 3: radial only
	GridSpec gs{};                // mode 1
	SearchPlan* gring = nullptr;  // mode 1: ring of the grid correlation
	int np = 0;                   // mode 1: micrograph columns padded to 32
	int nchunks = 0;              // mode 1: top-k row chunks
	SearchProgram pg{};           // mode 2: narrowed search program
	float window = 300.0f;        // mode 2: half-width of the defocus window (A)
	float conf_thr = 0.8f;        // mode 2: proposal confidence threshold
	float* radial_out = nullptr;  // mode 3: device [n][ld]
	int radial_ld = 0;
};

int record_spectra(gctf_ctx* c, const RescalePlan& rp, TileTables& tt, const void* imgs, int dtype, int n,
	long long stride)
{
	const int T = c->prm.tile_size;
	const TileGrid g = make_grid(rp.nnx, rp.nny, T);
	const int C = std::max(1, std::min(c->chunk, n));
	k_state_init<<<(n + 63) / 64, 64, 0, c->stream>>>(c->states.as<SearchState>(), n, c->prm.ext_phase_deg,
		std::max(c->prm.ext_phase_range_deg, 0.0f));
	const size_t esz = dtype == GCTF_DTYPE_F16 ? 2 : 4;
	for (int i0 = 0; i0 < n; i0 += C) {
		const int m = std::min(C, n - i0);
		const char* src = (const char*)imgs + (size_t)i0 * stride * esz;
		if (rp.active) {
			float* pad = c->padc.as<float>();
			const long long ps = (long long)rp.pitch * rp.nny;
			CHK(launch_rescale(c, rp, src, dtype, stride, m, pad, ps, &g));
			CHK(launch_spectrum<float>(c, g, tt, pad, rp.pitch, ps, m, i0, true));
		} else if (dtype == GCTF_DTYPE_F16) {
			CHK(launch_spectrum<_Float16>(c, g, tt, (const _Float16*)src, rp.nx, stride, m, i0));
		} else {
			CHK(launch_spectrum<float>(c, g, tt, (const float*)src, rp.nx, stride, m, i0));
		}
	}
	return launch_post(c, tt, n);
}

template <int NT>
void launch_gemm_t(gctf_ctx* c, const BatchJob& j, int kpad)
{
	const SearchPlan& r = *j.gring;
	const int per_wg = 32 * kGemmWaves;
	PROF(c, "k_grid_gemm", k_grid_gemm<NT><<<(j.gs.m + per_wg - 1) / per_wg, 64 * kGemmWaves, 0, c->stream>>>(
		j.gs, r.cc, c->g_gt.as<float>(), kpad / kGemmKs, c->g_bt.as<__bf16>(), c->g_fn.as<double>(),
		c->prm.ext_phase_deg, c->g_cc.as<float>()));
}

int launch_grid(gctf_ctx* c, const BatchJob& j, int n)
{
	hipStream_t s = c->stream;
	const SearchPlan& r = *j.gring;
	const int T = r.T, cx = T / 2 + 1;
	const int K = r.ring.n, kpad = (K + kGemmKs - 1) / kGemmKs * kGemmKs;
	PROF(c, "k_ring_gather", k_ring_gather_uniform<<<n, 1024, 0, s>>>(c->spect.as<float>(), T * cx, r.ring,
		c->ring_spec.as<float>(), c->ring_ss.as<double>()));
	PROF(c, "k_grid_tables", k_grid_tables<<<(kpad + 255) / 256, 256, 0, s>>>(r.cc, r.ring,
		c->prm.ext_phase_deg, kpad, c->g_gt.as<float>()));
	PROF(c, "k_grid_pack", k_grid_pack<<<j.np, 256, 0, s>>>(c->ring_spec.as<float>(), K, kpad, n, j.np,
		c->g_bt.as<__bf16>(), c->g_fn.as<double>()));
	switch (j.np / 32) {
	case 1: launch_gemm_t<1>(c, j, kpad); break;
	case 2: launch_gemm_t<2>(c, j, kpad); break;
	case 3: launch_gemm_t<3>(c, j, kpad); break;
	case 4: launch_gemm_t<4>(c, j, kpad); break;
	case 5: launch_gemm_t<5>(c, j, kpad); break;
	case 6: launch_gemm_t<6>(c, j, kpad); break;
	case 7: launch_gemm_t<7>(c, j, kpad); break;
	default: launch_gemm_t<8>(c, j, kpad); break;
	}
	const int rows = (j.gs.m + j.nchunks - 1) / j.nchunks;
	PROF(c, "k_grid_topk", k_grid_topk<<<j.nchunks, 256, 0, s>>>(c->g_cc.as<float>(), j.gs.m, j.np, n, rows,
		c->g_part.as<float2>()));
	PROF(c, "k_grid_merge", k_grid_merge<<<n, 256, 0, s>>>(c->g_part.as<float2>(), j.nchunks, j.np,
		c->g_top.as<int>()));
	PROF(c, "k_grid_rescore", k_grid_rescore<<<dim3(kGridTopK / 8, kPixChunks, n), 256, 0, s>>>(j.gs, r.cc,
		r.ring, c->ring_spec.as<float>(), c->g_top.as<int>(), c->prm.ext_phase_deg, c->g_rs.as<double2>()));
	PROF(c, "k_grid_pick", k_grid_pick<<<(n + 63) / 64, 64, 0, s>>>(j.gs, c->g_rs.as<double2>(),
		c->ring_ss.as<double>(), c->g_top.as<int>(), c->prm.ext_phase_deg, c->states.as<SearchState>(),
		c->g_rescored.as<float>(), n));
	HIPCHK(c, hipGetLastError());
	return GCTF_OK;
}

int record_batch(gctf_ctx* c, const RescalePlan& rp, SearchPlan& sp, TileTables& tt, const void* imgs,
	int dtype, int n, long long stride, const BatchJob& job)
{
	CHK(record_spectra(c, rp, tt, imgs, dtype, n, stride));
	if (job.mode == 1) return launch_grid(c, job, n);
	if (job.mode == 2 || job.mode == 3) {
		const int T = sp.T, cx = T / 2 + 1;
		hipStream_t s = c->stream;
		PROF(c, "k_radial", k_radial<<<dim3((cx + 63) / 64, n), 64, 0, s>>>(c->spect.as<float>(), cx, T, tt.roff.as<int>(),
			tt.rcnt.as<int>(), tt.ridx.as<int>(), c->radial.as<float>(), cx));
		if (job.mode == 3) {
			HIPCHK(c, hipMemcpy2DAsync(job.radial_out, sizeof(float) * job.radial_ld, c->radial.p,
				sizeof(float) * cx, sizeof(float) * std::min(cx, job.radial_ld), n, hipMemcpyDeviceToDevice, s));
			return GCTF_OK;
		}
		hipError_t ne = hipSuccess;
		PROF(c, "net_forward", ne = net_forward(c->netb, c->net_params.as<float>(), c->radial.as<float>(), cx, n,
			true, c->net_out.as<float>(), s));
		HIPCHK(c, ne);
		PROF(c, "k_proposal", k_proposal<<<(n + 63) / 64, 64, 0, s>>>(c->net_out.as<float>(), n, job.window,
			job.conf_thr, c->props.as<float>(), c->refine.as<RefineArgs>(), c->states.as<SearchState>()));
		CHK(fill_plans(c, 0, n, sp));
		CHK(run_2d_program(c, view_2d(c, T, 0), n, NX_BEGIN_PROPOSAL, job.pg,
			std::max(c->prm.ext_phase_range_deg, 0.0f)));
		PROF(c, "k_learned_check", k_learned_check<<<(n + 63) / 64, 64, 0, s>>>(c->states.as<SearchState>(), n,
			job.conf_thr));
		return GCTF_OK;
	}
	CHK(fill_plans(c, 0, n, sp));
	CHK(launch_1d(c, tt, 0, n, 0, 0.f, 0.f, c->prm.ext_phase_range_deg > 0.0f));
	CHK(launch_2d(c, sp.T, 0, n, NX_BEGIN_DO2D, std::max(c->prm.ext_phase_range_deg, 0.0f)));
	return GCTF_OK;
}

std::string shape_key(int nx, int ny, int dtype, float pix)
{
	char k[96];
	snprintf(k, sizeof k, "%d/%d/%d/%a", nx, ny, dtype, pix);
	return k;
}

// Chunk and tile-split counts chosen by gctf_autotune for this shape, in
// force for the lifetime of the scope only: the context's own values (the
// library defaults or gctf_tune's) come back when the call returns, so one
// shape's tuning never reaches a call on another shape.
struct TuneScope {
	gctf_ctx* c;
	int chunk, splits, cpb;
	TuneScope(gctf_ctx* ctx, int nx, int ny, int dtype, float pix)
		: c(ctx), chunk(ctx->chunk), splits(ctx->tile_splits), cpb(ctx->cpb)
	{
		auto it = ctx->tuned.find(shape_key(nx, ny, dtype, pix));
		if (it == ctx->tuned.end()) return;
		ctx->chunk = it->second.chunk;
		ctx->tile_splits = it->second.splits;
		ctx->cpb = it->second.cpb;
	}
	~TuneScope() { c->chunk = chunk; c->tile_splits = splits; c->cpb = cpb; }
	TuneScope(const TuneScope&) = delete;
	TuneScope& operator=(const TuneScope&) = delete;
};

// Record the launches of `record` into an executable graph cached under
// `key` (re-captured after any device buffer moved) and launch it; without
// graphs (or while profiling) the launches go straight to the stream.
template <class F>
int run_graph(gctf_ctx* ctx, const std::string& key, F&& record)
{
	if (!ctx->use_graph || ctx->prof) return record();
	GraphEntry& ge = ctx->graphs[key];
	if (!ge.exec || ge.gen != ctx->gen) {
		if (ge.exec) { (void)hipGraphExecDestroy(ge.exec); ge.exec = nullptr; }
		hipGraph_t graph = nullptr;
		HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
		int r = record();
		hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
		if (r != GCTF_OK) { if (graph) (void)hipGraphDestroy(graph); return r; }
		HIPCHK(ctx, e);
		hipError_t ei = hipGraphInstantiate(&ge.exec, graph, nullptr, nullptr, 0);
		(void)hipGraphDestroy(graph);
		HIPCHK(ctx, ei);
		ge.gen = ctx->gen;
	}
	HIPCHK(ctx, hipGraphLaunch(ge.exec, ctx->stream));
	return GCTF_OK;
}

// Shared driver of the batched entry points: plans, buffers, graph capture
// (one executable graph per input pointer, shape and job) and read-back.
int run_batch(gctf_ctx* ctx, const void* dev_imgs, int dtype, int n, int nx, int ny, long long stride,
	float pix, gctf_result* out, BatchJob& job, const char* job_key)
{
	if (!dev_imgs || !out || n < 1) return fail(ctx, GCTF_ERR_ARG, "null input/output or n < 1");
	if (dtype != GCTF_DTYPE_F32 && dtype != GCTF_DTYPE_F16) return fail(ctx, GCTF_ERR_ARG, "dtype");
	if (stride < (long long)nx * ny) return fail(ctx, GCTF_ERR_ARG, "stride smaller than one image");
	HIPCHK(ctx, hipSetDevice(ctx->device));
	TuneScope tune(ctx, nx, ny, dtype, pix);
	RescalePlan rp;
	CHK(make_rescale_plan(ctx, nx, ny, pix, rp));
	const int T = ctx->prm.tile_size;
	if (rp.nnx < T || rp.nny < T) return fail(ctx, GCTF_ERR_ARG, "image smaller than one tile");
	const TileGrid g = make_grid(rp.nnx, rp.nny, T);
	if (g.ntx < 1 || g.nty < 1) return fail(ctx, GCTF_ERR_ARG, "no tile fits the image");
	SearchPlan* sp;
	TileTables* tt;
	CHK(get_search_plan(ctx, ctx->prm.kv, ctx->prm.cs_mm, ctx->prm.amp_contrast, rp.pixn, T, &sp));
	CHK(get_tile_tables(ctx, T, &tt));
	int ring_n = sp->ring.n;
	ctx->kstride = sp->ring.n;
	if (job.mode == 1) ring_n = std::max(ring_n, job.gring->ring.n);
	const int C = std::max(1, std::min(ctx->chunk, n));
	CHK(ensure_image_buffers(ctx, n, T, ring_n));
	CHK(ensure_spectrum_buffers(ctx, g, T, C));
	if (rp.active) {
		CHK(ensure_rescale_buffers(ctx, rp, C));
		CHK(ensure_buf(ctx, ctx->padc, sizeof(float) * (size_t)rp.pitch * rp.nny * C));
	}
	if (job.mode == 2) {
		if (!ctx->have_net) return fail(ctx, GCTF_ERR_STATE, "no proposal-network parameters installed");
		if (!ctx->netb || net_buffers_capacity(ctx->netb) < n) {
			net_buffers_destroy(ctx->netb);
			ctx->netb = nullptr;
			HIPCHK(ctx, net_buffers_create(n, false, &ctx->netb));
			ctx->gen++;
		}
		CHK(ensure_buf(ctx, ctx->net_out, sizeof(float) * kNetOut * n));
		CHK(ensure_buf(ctx, ctx->props, sizeof(float) * 4 * n));
	}
	char key[320];
	snprintf(key, sizeof key, "%p/%d/%d/%d/%d/%lld/%a/%d/%d/%d/%s", dev_imgs, dtype, n, nx, ny, stride, pix, C,
		ctx->tile_splits, ctx->cpb, job_key);
	CHK(run_graph(ctx, key, [&]() { return record_batch(ctx, rp, *sp, *tt, dev_imgs, dtype, n, stride, job); }));
	return read_results(ctx, n, out);
}

}  // namespace

int gctf_estimate_batch(gctf_ctx* ctx, const void* dev_imgs, int dtype, int n, int nx, int ny,
	long long stride, float pix, gctf_result* out)
{
	if (!ctx) return GCTF_ERR_ARG;
	BatchJob job;
	return run_batch(ctx, dev_imgs, dtype, n, nx, ny, stride, pix, out, job, "plain");
}

int gctf_grid_search_batch(gctf_ctx* ctx, const void* dev_imgs, int dtype, int n, int nx, int ny,
	long long stride, float pix, const gctf_grid* grid, gctf_result* out, float* cc_out, int* top_out,
	float* rescored_out)
{
	if (!ctx) return GCTF_ERR_ARG;
	if (!grid || grid->n_df < 1 || grid->n_ast < 1 || grid->n_az < 1 || !(grid->res_lo > grid->res_hi)
		|| !(grid->res_hi > 0))
		return fail(ctx, GCTF_ERR_ARG, "bad grid");
	if (n > 256) return fail(ctx, GCTF_ERR_UNSUPPORTED, "grid search takes at most 256 micrographs per call");
	const long long m = (long long)grid->n_df * grid->n_ast * grid->n_az;
	if (m > (1LL << 26)) return fail(ctx, GCTF_ERR_UNSUPPORTED, "candidate grid too large");
	BatchJob job;
	job.mode = 1;
	GridSpec& g = job.gs;
	g.n_df = grid->n_df; g.n_ast = grid->n_ast; g.n_az = grid->n_az;
	g.df_lo = grid->df_lo;
	g.df_step = grid->n_df > 1 ? (grid->df_hi - grid->df_lo) / (float)(grid->n_df - 1) : 0.0f;
	g.ast_lo = grid->ast_lo;
	g.ast_step = grid->n_ast > 1 ? (grid->ast_hi - grid->ast_lo) / (float)(grid->n_ast - 1) : 0.0f;
	g.az_lo = grid->az_lo;
	g.az_step = (grid->az_hi - grid->az_lo) / (float)grid->n_az;
	g.m = (int)m;
	job.np = (n + 31) / 32 * 32;
	job.nchunks = (int)std::min<long long>(4096, (m + 255) / 256);   // about 256 rows per top-k thread
	// the ring plan of the grid correlation (resolution bounds in A)
	HIPCHK(ctx, hipSetDevice(ctx->device));
	RescalePlan rp;
	CHK(make_rescale_plan(ctx, nx, ny, pix, rp));
	CHK(get_search_plan(ctx, ctx->prm.kv, ctx->prm.cs_mm, ctx->prm.amp_contrast, rp.pixn, ctx->prm.tile_size,
		&job.gring, grid->res_lo, grid->res_hi));
	const int K = job.gring->ring.n, kpad = (K + kGemmKs - 1) / kGemmKs * kGemmKs;
	CHK(ensure_buf(ctx, ctx->g_bt, sizeof(__bf16) * (size_t)kpad * job.np));
	CHK(ensure_buf(ctx, ctx->g_gt, sizeof(float) * (size_t)kpad * kGemmTabs));
	CHK(ensure_buf(ctx, ctx->g_fn, sizeof(double) * job.np));
	CHK(ensure_buf(ctx, ctx->g_cc, sizeof(float) * (size_t)g.m * job.np));
	CHK(ensure_buf(ctx, ctx->g_part, sizeof(float2) * (size_t)job.nchunks * job.np * kGridTopK));
	CHK(ensure_buf(ctx, ctx->g_top, sizeof(int) * (size_t)n * kGridTopK));
	CHK(ensure_buf(ctx, ctx->g_rs, sizeof(double2) * (size_t)n * kGridTopK * kPixChunks));
	CHK(ensure_buf(ctx, ctx->g_rescored, sizeof(float) * (size_t)n * kGridTopK));
	char key[160];
	snprintf(key, sizeof key, "grid/%d/%d/%d/%a/%a/%a/%a/%a/%a/%a/%a", g.n_df, g.n_ast, g.n_az, g.df_lo,
		g.df_step, g.ast_lo, g.ast_step, g.az_lo, g.az_step, grid->res_lo, grid->res_hi);
	CHK(run_batch(ctx, dev_imgs, dtype, n, nx, ny, stride, pix, out, job, key));
	if (cc_out) {
		std::vector<float> all((size_t)g.m * job.np);
		HIPCHK(ctx, hipMemcpy(all.data(), ctx->g_cc.p, sizeof(float) * all.size(), hipMemcpyDeviceToHost));
		for (int i = 0; i < n; ++i)
			for (int mm = 0; mm < g.m; ++mm) cc_out[(size_t)i * g.m + mm] = all[(size_t)mm * job.np + i];
	}
	if (top_out)
		HIPCHK(ctx, hipMemcpy(top_out, ctx->g_top.p, sizeof(int) * n * kGridTopK, hipMemcpyDeviceToHost));
	if (rescored_out)
		HIPCHK(ctx, hipMemcpy(rescored_out, ctx->g_rescored.p, sizeof(float) * n * kGridTopK,
			hipMemcpyDeviceToHost));
	return GCTF_OK;
}



int gctf_radial_batch(gctf_ctx* ctx, const void* dev_imgs, int dtype, int n, int nx, int ny, long long stride,
	float pix, float* d_radial_out, int ld)
{
	if (!ctx) return GCTF_ERR_ARG;
	if (!d_radial_out || ld < 1) return fail(ctx, GCTF_ERR_ARG, "null radial output or ld < 1");
	if (n < 1) return fail(ctx, GCTF_ERR_ARG, "n < 1");
	BatchJob job;
	job.mode = 3;
	job.radial_out = d_radial_out;
	job.radial_ld = ld;
	std::vector<gctf_result> tmp((size_t)n);
	char key[96];
	snprintf(key, sizeof key, "radial/%p/%d", (void*)d_radial_out, ld);
	return run_batch(ctx, dev_imgs, dtype, n, nx, ny, stride, pix, tmp.data(), job, key);
}

int gctf_set_proposal_params(gctf_ctx* ctx, const float* host_params, long long count)
{
	if (!ctx) return GCTF_ERR_ARG;
	if (!host_params || count != kNetParams) return fail(ctx, GCTF_ERR_ARG, "proposal parameter count");
	HIPCHK(ctx, hipSetDevice(ctx->device));
	CHK(ensure_buf(ctx, ctx->net_params, sizeof(float) * kNetParams));
	HIPCHK(ctx, hipMemcpy(ctx->net_params.p, host_params, sizeof(float) * kNetParams, hipMemcpyHostToDevice));
	ctx->have_net = true;
	return GCTF_OK;
}

int gctf_estimate_batch_learned(gctf_ctx* ctx, const void* dev_imgs, int dtype, int n, int nx, int ny,
	long long stride, float pix, const gctf_search_program* prog, float conf_threshold, float window_A,
	gctf_result* out, gctf_proposal* prop_out)
{
	if (!ctx) return GCTF_ERR_ARG;
	BatchJob job;
	job.mode = 2;
	SearchProgram& pg = job.pg;
	if (prog) {
		pg.n_ast = prog->n_ast; pg.n_ang = prog->n_ang; pg.n_mag = prog->n_mag; pg.n_angl = prog->n_angl;
		pg.n_df = prog->n_df; pg.iters = prog->iters; pg.df_minstep = prog->df_minstep;
		pg.fine_n = prog->fine_n; pg.fine_step = prog->fine_step;
	} else {
		gctf_search_program d;
		gctf_learned_program_default(&d);
		pg.n_ast = d.n_ast; pg.n_ang = d.n_ang; pg.n_mag = d.n_mag; pg.n_angl = d.n_angl;
		pg.n_df = d.n_df; pg.iters = d.iters; pg.df_minstep = d.df_minstep;
		pg.fine_n = d.fine_n; pg.fine_step = d.fine_step;
	}
	if (pg.n_ast < 1 || pg.n_ang < 1 || pg.n_ast * pg.n_ang > kMaxCands || pg.n_mag < 2 || pg.n_angl < 2
		|| pg.n_df < 2 || pg.n_mag >= kMaxCands || pg.n_angl >= kMaxCands || pg.n_df >= kMaxCands
		|| pg.iters < 2 || pg.iters > 200 || !(pg.df_minstep > 0) || pg.fine_n < 0 || pg.fine_n >= kMaxCands
		|| (pg.fine_n >= 2 && !(pg.fine_step > 0)))
		return fail(ctx, GCTF_ERR_ARG, "search program outside the supported counts");
	if (!(window_A > 0)) return fail(ctx, GCTF_ERR_ARG, "window must be positive");
	job.window = window_A;
	job.conf_thr = conf_threshold;
	char key[192];
	snprintf(key, sizeof key, "learned/%d/%d/%d/%d/%d/%d/%a/%d/%a/%a/%a", pg.n_ast, pg.n_ang, pg.n_mag,
		pg.n_angl, pg.n_df, pg.iters, pg.df_minstep, pg.fine_n, pg.fine_step, window_A, conf_threshold);
	CHK(run_batch(ctx, dev_imgs, dtype, n, nx, ny, stride, pix, out, job, key));
	// fallback pass: the plain search on the micrographs the learned path did not settle
	std::vector<int> fb((size_t)n);
	int nfb = 0;
	for (int i = 0; i < n; ++i) { fb[i] = ctx->h_pinned[i].fallback; nfb += fb[i]; }
	if (nfb > 0) {
		RescalePlan rp;
		CHK(make_rescale_plan(ctx, nx, ny, pix, rp));
		SearchPlan* sp;
		TileTables* tt;
		CHK(get_search_plan(ctx, ctx->prm.kv, ctx->prm.cs_mm, ctx->prm.amp_contrast, rp.pixn,
			ctx->prm.tile_size, &sp));
		CHK(get_tile_tables(ctx, ctx->prm.tile_size, &tt));
		const float range = std::max(ctx->prm.ext_phase_range_deg, 0.0f);
		SearchState* st = ctx->states.as<SearchState>();
		// one graph for every fallback pattern: which micrographs sit out is data
		TuneScope tune(ctx, nx, ny, dtype, pix);
		// few micrographs take part: smaller candidate blocks put more
		// workgroups on the chip (the results do not depend on the blocking)
		if (4 * nfb <= n) ctx->cpb = 2;
		else if (2 * nfb <= n) ctx->cpb = std::min(ctx->cpb, 4);
		char fkey[160];
		snprintf(fkey, sizeof fkey, "fallback/%d/%d/%a/%a/%a/%d", n, ctx->prm.tile_size, sp->pix, range,
			ctx->prm.ext_phase_deg, ctx->cpb);
		CHK(run_graph(ctx, fkey, [&]() -> int {
			k_fallback_prepare<<<(n + 63) / 64, 64, 0, ctx->stream>>>(st, n, ctx->prm.ext_phase_deg, range);
			CHK(fill_plans(ctx, 0, n, *sp));
			CHK(launch_1d(ctx, *tt, 0, n, 0, 0.f, 0.f, range > 0.0f));
			CHK(launch_2d(ctx, ctx->prm.tile_size, 0, n, NX_BEGIN_DO2D, range));
			k_clear_skip<<<(n + 63) / 64, 64, 0, ctx->stream>>>(st, n);
			HIPCHK(ctx, hipGetLastError());
			return GCTF_OK;
		}));
		CHK(read_results(ctx, n, out));
	}
	if (prop_out) {
		std::vector<float> pr((size_t)4 * n);
		HIPCHK(ctx, hipMemcpy(pr.data(), ctx->props.p, sizeof(float) * 4 * n, hipMemcpyDeviceToHost));
		for (int i = 0; i < n; ++i) {
			prop_out[i].df_mean = pr[4 * i];
			prop_out[i].astig = pr[4 * i + 1];
			prop_out[i].azimuth_deg = pr[4 * i + 2];
			prop_out[i].confidence = pr[4 * i + 3];
			prop_out[i].fallback = fb[i];
		}
	}
	return GCTF_OK;
}

int gctf_learned_program_default(gctf_search_program* out)
{
	if (!out) return GCTF_ERR_ARG;
	out->n_ast = 11; out->n_ang = 18;
	out->n_mag = 11; out->n_angl = 27;
	out->n_df = 11;
	out->iters = 6;
	out->df_minstep = 20.0f;
	out->fine_n = 9;
	out->fine_step = 5.0f;
	return GCTF_OK;
}

// ---- tilt series (config: one series, same shape) --------------------------
namespace {

struct SeriesCtl {
	int ref;    // CFindSeriesCtfs's reference package: best score so far
	int next;   // packages order[0 .. next) are settled
	int full;   // package that needs the full search this round, or -1
	int pad;
};

// Packages in processing order (CLoadImages.cpp:71-91): the zero-tilt
// package z first, then the others in file order.  The first package's full
// search has run, so it is the reference and settled.
__global__ void k_series_init(SeriesCtl* __restrict__ ctl, int* __restrict__ order, int n, int z)
{
	for (int j = threadIdx.x; j < n; j += blockDim.x) order[j] = j == 0 ? z : (j - 1 < z ? j - 1 : j);
	if (threadIdx.x == 0) { ctl->ref = z; ctl->next = 1; ctl->full = -1; }
}

// CFindSeriesCtfs::mProcessRefine arguments (CFindSeriesCtfs.cpp:168-203)
// from the reference package's results, for every package not yet settled;
// settled packages sit the round out.  Every pending package refines around
// the same reference: the results of those after the next change of
// reference are discarded by k_series_walk and computed again.
__global__ void k_series_round_prep(SearchState* __restrict__ st, RefineArgs* __restrict__ refine,
	const SeriesCtl* __restrict__ ctl, const int* __restrict__ order, int n, float pix_in, int tomo,
	int phase_on)
{
	const SearchState& R = st[ctl->ref];
	RefineArgs r;
	const float range = 5000.0f * pix_in * pix_in;
	r.dfmean[0] = 0.5f * (R.df_min + R.df_max);
	r.dfmean[1] = fmaxf(r.dfmean[0] * 0.5f, range);
	// CFindCtfHelp::CalcAstRatio (CFindCtfHelp.cpp:5-11)
	const float diff = (R.df_max - R.df_min) * 0.5f, mean = (R.df_max + R.df_min) * 0.5f;
	r.ast[0] = mean <= 0.0f ? 0.0f : diff / mean;
	r.ang[0] = R.ast_ang;
	r.phase[0] = R.ext_phase;
	r.ast[1] = 0.0f; r.ang[1] = 0.0f; r.phase[1] = 0.0f;
	if (!tomo) {
		r.ast[1] = 0.04f;
		r.ang[1] = 30.0f;
		if (phase_on) r.phase[1] = 40.0f;
	}
	const int next = ctl->next;
	__syncthreads();   // every thread has read the reference before any state changes
	for (int j = threadIdx.x; j < n; j += blockDim.x) {
		const int k = order[j];
		if (j < next) {
			st[k].skip = 1;
		} else {
			st[k].skip = 0;
			refine[k] = r;
		}
	}
}

// The sequential package rules over the round's refines, in processing
// order: a better score makes the package the reference (:200-202), a score
// <= 0 sends it to the full search (:144), which makes it the reference too
// (:164).  Either ends the round after that package: the later ones refined
// around a reference that is no longer current.  One thread.
__global__ void k_series_walk(SearchState* __restrict__ st, SeriesCtl* __restrict__ ctl,
	const int* __restrict__ order, int n, float p0, float range)
{
	int ref = ctl->ref, j = ctl->next, full = -1;
	for (; j < n; ++j) {
		const int k = order[j];
		const SearchState& s = st[k];
		bool stop = false;
		if (st[ref].score < s.score) { ref = k; stop = true; }
		if (s.score <= 0.0f) { full = k; stop = true; }
		if (stop) { ++j; break; }
	}
	ctl->ref = ref;
	ctl->next = j;
	ctl->full = full;
	// only the package that needs it takes part in the full search, with
	// CFindCtfBase::SetPhase's values
	for (int i = 0; i < n; ++i) {
		SearchState& s = st[i];
		s.skip = i == full ? 0 : 1;
		if (i == full) {
			s.fallback = 1;
			s.ext_phase = p0;
			s.phase_range = range;
		}
	}
}

// mProcessFull makes its package the reference (:164).
__global__ void k_series_after_full(SearchState* __restrict__ st, SeriesCtl* __restrict__ ctl, int n)
{
	if (ctl->full >= 0) ctl->ref = ctl->full;
	for (int i = 0; i < n; ++i) st[i].skip = 0;
}

// CInputFolder::FindZeroTilt (CInputFolder.cpp:49-68)
int zero_tilt(const float* tilts, int n)
{
	if (n < 2) return 0;
	if (fabsf(tilts[0] - tilts[1]) < 0.0001f) return n / 2;
	int z = -1;
	float m = 10000.0f;
	for (int i = 0; i < n; ++i) {
		const float t = fabsf(tilts[i]);
		if (t >= m) continue;
		m = t;
		z = i;
	}
	return z < 0 ? n / 2 : z;
}

}  // namespace

int gctf_estimate_series(gctf_ctx* ctx, const void* dev_imgs, int dtype, int n, int nx, int ny,
	long long stride, float pix, const float* tilts_deg, int tomo, gctf_result* out)
{
	if (!ctx) return GCTF_ERR_ARG;
	if (!dev_imgs || !out || !tilts_deg || n < 1) return fail(ctx, GCTF_ERR_ARG, "null input/output or n < 1");
	if (dtype != GCTF_DTYPE_F32 && dtype != GCTF_DTYPE_F16) return fail(ctx, GCTF_ERR_ARG, "dtype");
	if (stride < (long long)nx * ny) return fail(ctx, GCTF_ERR_ARG, "stride smaller than one image");
	HIPCHK(ctx, hipSetDevice(ctx->device));
	TuneScope tune(ctx, nx, ny, dtype, pix);
	RescalePlan rp;
	CHK(make_rescale_plan(ctx, nx, ny, pix, rp));
	const int T = ctx->prm.tile_size;
	if (rp.nnx < T || rp.nny < T) return fail(ctx, GCTF_ERR_ARG, "image smaller than one tile");
	const TileGrid g = make_grid(rp.nnx, rp.nny, T);
	if (g.ntx < 1 || g.nty < 1) return fail(ctx, GCTF_ERR_ARG, "no tile fits the image");
	SearchPlan* sp;
	TileTables* tt;
	CHK(get_search_plan(ctx, ctx->prm.kv, ctx->prm.cs_mm, ctx->prm.amp_contrast, rp.pixn, T, &sp));
	CHK(get_tile_tables(ctx, T, &tt));
	const int C = std::max(1, std::min(ctx->chunk, n));
	CHK(ensure_image_buffers(ctx, n, T, sp->ring.n));
	ctx->kstride = sp->ring.n;
	CHK(ensure_spectrum_buffers(ctx, g, T, C));
	if (rp.active) {
		CHK(ensure_rescale_buffers(ctx, rp, C));
		CHK(ensure_buf(ctx, ctx->padc, sizeof(float) * (size_t)rp.pitch * rp.nny * C));
	}
	CHK(ensure_buf(ctx, ctx->series, sizeof(SeriesCtl) + sizeof(int) * (size_t)n));
	// CLoadImages.cpp:71-91: the zero-tilt package first, the others in order
	const int z = tomo ? zero_tilt(tilts_deg, n) : 0;
	const float p0 = ctx->prm.ext_phase_deg, range = std::max(ctx->prm.ext_phase_range_deg, 0.0f);
	const int phase_on = ctx->prm.ext_phase_range_deg > 0.0f ? 1 : 0;
	const float refine_phase = (!tomo && phase_on) ? 40.0f : 0.0f;
	SeriesCtl* ctl = ctx->series.as<SeriesCtl>();
	int* order = reinterpret_cast<int*>(ctl + 1);
	SearchState* st = ctx->states.as<SearchState>();
	// the spectra, then mProcessFull for the first package (:155-166)
	auto record_first = [&]() -> int {
		CHK(record_spectra(ctx, rp, *tt, dev_imgs, dtype, n, stride));
		CHK(fill_plans(ctx, 0, n, *sp));
		CHK(launch_1d(ctx, *tt, z, 1, 0, 0.f, 0.f, ctx->prm.ext_phase_range_deg > 0.0f));
		CHK(launch_2d(ctx, T, z, 1, NX_BEGIN_DO2D, range));
		k_series_init<<<1, 256, 0, ctx->stream>>>(ctl, order, n, z);
		HIPCHK(ctx, hipGetLastError());
		return GCTF_OK;
	};
	// one round: every pending package refines around the current reference
	// in one batch, and the package rules settle them in order up to the
	// first change of reference or failed refine
	auto record_round = [&]() -> int {
		k_series_round_prep<<<1, 256, 0, ctx->stream>>>(st, ctx->refine.as<RefineArgs>(), ctl, order, n, pix,
			tomo, phase_on);
		CHK(launch_2d(ctx, T, 0, n, NX_BEGIN_REFINE, refine_phase));
		k_series_walk<<<1, 1, 0, ctx->stream>>>(st, ctl, order, n, p0, range);
		HIPCHK(ctx, hipGetLastError());
		return GCTF_OK;
	};
	// the full search of the package whose refine failed (the others sit out)
	auto record_full = [&]() -> int {
		CHK(launch_1d(ctx, *tt, 0, n, 0, 0.f, 0.f, ctx->prm.ext_phase_range_deg > 0.0f));
		CHK(launch_2d(ctx, T, 0, n, NX_BEGIN_DO2D, range));
		k_series_after_full<<<1, 1, 0, ctx->stream>>>(st, ctl, n);
		HIPCHK(ctx, hipGetLastError());
		return GCTF_OK;
	};
	char key[192];
	snprintf(key, sizeof key, "series/%p/%d/%d/%d/%d/%lld/%a/%d/%d/%d/%d/%d/%a/%a", dev_imgs, dtype, n, nx, ny,
		stride, pix, C, ctx->tile_splits, ctx->cpb, tomo, z, p0, range);
	CHK(run_graph(ctx, key, record_first));
	char rkey[192], fkey[192];
	snprintf(rkey, sizeof rkey, "series-round/%d/%a/%d/%d/%d/%a/%a", n, pix, ctx->cpb, tomo, phase_on, p0, range);
	snprintf(fkey, sizeof fkey, "series-full/%d/%a/%d/%a/%a", n, pix, ctx->cpb, p0, range);
	// typically one round: the zero-tilt package keeps the best score; each
	// change of reference or full search costs one more, at most n - 1
	for (int round = 1; round < n; ++round) {
		CHK(run_graph(ctx, rkey, record_round));
		SeriesCtl h{};
		HIPCHK(ctx, hipMemcpyAsync(&h, ctl, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
		HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
		if (h.full >= 0) CHK(run_graph(ctx, fkey, record_full));
		if (h.next >= n) break;
	}
	// no package sits out later calls on this context
	k_clear_skip<<<(n + 63) / 64, 64, 0, ctx->stream>>>(st, n);
	HIPCHK(ctx, hipGetLastError());
	return read_results(ctx, n, out);
}

// ---- uneven batch ------------------------------------------------------------
// The micrographs of a list run in groups of consecutive jobs read with the
// same pixel type whose row-pass output fits kListRtBytes; buffer offsets
// restart at every group.  In a group, each rescale stage is one launch per
// kernel class over all its micrographs (k_*_list), and the tile passes are
// one launch each over the group's tile work list.  The row pass runs as
// persistent workgroups pulling items from a work queue (k_tile_rows_queue),
// or one workgroup per item (k_tile_rows_list) with gctf_tune's list mode
// -1; the column pass runs one workgroup per item (k_tile_cols_list).
constexpr size_t kListRtBytes = 8ull << 30;

struct ListGroup {
	int j0 = 0, nj = 0;      // jobs [j0, j0 + nj)
	int f16 = 0;             // pixels read as float16
	int items = 0;           // row-pass items
	int max_nr = 0;          // largest block-row count among jobs that need row sums
	int max_np = 0, max_split = 0;
	size_t pad = 0, rowsum = 0, sums = 0, rt = 0, part = 0, cbuf = 0;   // elements
	int q = 0;               // its row pass's work-queue counter
};

// One rescale stage's launch over the micrographs of one kernel class.
enum RsStage { RS_R2C = 0, RS_TRANSPOSE = 1, RS_COLS = 2, RS_C2R = 3 };
struct RsClass {
	int stage = 0, variant = 0, maxn = 0, f16 = 0;   // variant: half rows (R2C, C2R) or split columns
	std::vector<int> entries, first;                 // rescale jobs and their first items
	int items = 0;
	size_t first_off = 0, job_off = 0;               // offsets into the uploaded int table
};

int pow2_class(int m) { return m <= 2048 ? 2048 : m <= 4096 ? 4096 : m <= 8192 ? 8192 : 16384; }

// Class keys and workgroup items of the four stages of one micrograph.
int rescale_classes(gctf_ctx* c, const RescalePlan& rp, int f16, int key[4][3], int items[4])
{
	const int ncx = rp.nnx / 2 + 1;
	if (rp.rows_half && rp.rows->m > rp.rows_maxn) return fail(c, GCTF_ERR_UNSUPPORTED, "row length");
	if (rp.cols_split && std::max(rp.cf->m, rp.ci->m) > rp.cols_maxn)
		return fail(c, GCTF_ERR_UNSUPPORTED, "column length");
	if (rp.c2r_half && rp.c2r->m > rp.c2r_maxn) return fail(c, GCTF_ERR_UNSUPPORTED, "row length");
	key[RS_R2C][0] = rp.rows_half; key[RS_R2C][1] = rp.rows_half ? rp.rows_maxn : pow2_class(rp.rows->m);
	key[RS_R2C][2] = f16;
	items[RS_R2C] = rp.rows_half ? rp.ny : (rp.ny + 1) / 2;
	key[RS_TRANSPOSE][0] = 0; key[RS_TRANSPOSE][1] = 0; key[RS_TRANSPOSE][2] = 0;
	items[RS_TRANSPOSE] = ((ncx + 31) / 32) * ((rp.ny + 31) / 32);
	key[RS_COLS][0] = rp.cols_split;
	key[RS_COLS][1] = rp.cols_split ? rp.cols_maxn : pow2_class(std::max(rp.cf->m, rp.ci->m));
	key[RS_COLS][2] = 0;
	items[RS_COLS] = 8 * ((ncx + 7) / 8);                // a multiple of 8: xcd_group sees the XCD
	key[RS_C2R][0] = rp.c2r_half; key[RS_C2R][1] = rp.c2r_half ? rp.c2r_maxn : pow2_class(rp.c2r->m);
	key[RS_C2R][2] = 0;
	items[RS_C2R] = rp.c2r_half ? rp.nny : (rp.nny + 1) / 2;
	return GCTF_OK;
}

#define RS_LAUNCH(name, kern, NT)                                                                   \
	PROF(c, name, kern<<<k.items, NT, 0, s>>>(jobs, first, job, n))

template <typename Tin>
int launch_rs_r2c(gctf_ctx* c, const RsClass& k, const RescaleJob* jobs, const int* first, const int* job)
{
	hipStream_t s = c->stream;
	const int n = (int)k.entries.size();
	if (k.variant) {
		if (k.maxn == 8192) RS_LAUNCH("k_rescale_rows_r2c_list", (k_rescale_rows_r2c_list<Tin, 1024, 8192, true>), 1024);
		else RS_LAUNCH("k_rescale_rows_r2c_list", (k_rescale_rows_r2c_list<Tin, 1024, 16384, true>), 1024);
	} else if (k.maxn == 2048) RS_LAUNCH("k_rescale_rows_r2c_list", (k_rescale_rows_r2c_list<Tin, 256, 2048, false>), 256);
	else if (k.maxn == 4096) RS_LAUNCH("k_rescale_rows_r2c_list", (k_rescale_rows_r2c_list<Tin, 512, 4096, false>), 512);
	else if (k.maxn == 8192) RS_LAUNCH("k_rescale_rows_r2c_list", (k_rescale_rows_r2c_list<Tin, 1024, 8192, false>), 1024);
	else RS_LAUNCH("k_rescale_rows_r2c_list", (k_rescale_rows_r2c_list<Tin, 1024, 16384, false>), 1024);
	return GCTF_OK;
}

int launch_rs_class(gctf_ctx* c, const RsClass& k, const RescaleJob* jobs, const int* tab)
{
	hipStream_t s = c->stream;
	const int* first = tab + k.first_off;
	const int* job = tab + k.job_off;
	const int n = (int)k.entries.size();
	switch (k.stage) {
	case RS_R2C:
		CHK(k.f16 ? launch_rs_r2c<_Float16>(c, k, jobs, first, job) : launch_rs_r2c<float>(c, k, jobs, first, job));
		break;
	case RS_TRANSPOSE:
		RS_LAUNCH("k_transpose_list", k_transpose_list, 256);
		break;
	case RS_COLS:
		if (k.variant) {
			if (k.maxn == 8192) RS_LAUNCH("k_rescale_cols_list", (k_rescale_cols_list<1024, 8192, true>), 1024);
			else RS_LAUNCH("k_rescale_cols_list", (k_rescale_cols_list<1024, 16384, true>), 1024);
		} else if (k.maxn == 2048) RS_LAUNCH("k_rescale_cols_list", (k_rescale_cols_list<256, 2048, false>), 256);
		else if (k.maxn == 4096) RS_LAUNCH("k_rescale_cols_list", (k_rescale_cols_list<512, 4096, false>), 512);
		else if (k.maxn == 8192) RS_LAUNCH("k_rescale_cols_list", (k_rescale_cols_list<1024, 8192, false>), 1024);
		else RS_LAUNCH("k_rescale_cols_list", (k_rescale_cols_list<1024, 16384, false>), 1024);
		break;
	default:
		if (k.variant) {
			if (k.maxn == 8192) RS_LAUNCH("k_rescale_rows_c2r_list", (k_rescale_rows_c2r_list<1024, 8192, true>), 1024);
			else RS_LAUNCH("k_rescale_rows_c2r_list", (k_rescale_rows_c2r_list<1024, 16384, true>), 1024);
		} else if (k.maxn == 2048) RS_LAUNCH("k_rescale_rows_c2r_list", (k_rescale_rows_c2r_list<256, 2048, false>), 256);
		else if (k.maxn == 4096) RS_LAUNCH("k_rescale_rows_c2r_list", (k_rescale_rows_c2r_list<512, 4096, false>), 512);
		else if (k.maxn == 8192) RS_LAUNCH("k_rescale_rows_c2r_list", (k_rescale_rows_c2r_list<1024, 8192, false>), 1024);
		else RS_LAUNCH("k_rescale_rows_c2r_list", (k_rescale_rows_c2r_list<1024, 16384, false>), 1024);
		break;
	}
	HIPCHK(c, hipGetLastError());
	return GCTF_OK;
}
#undef RS_LAUNCH

// Workgroups of a persistent launch: as many as stay resident on the device
// at once (occupancy x compute units), at most `items`.
template <class K>
int resident_grid(gctf_ctx* c, K kernel, int threads, int items, int* out)
{
	int per_cu = 0, cus = 0;
	HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0));
	HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
	*out = std::max(1, std::min(items, std::max(1, per_cu) * std::max(1, cus)));
	return GCTF_OK;
}

template <typename Tin, int T>
int launch_tile_list_t(gctf_ctx* c, const TileTables& tt, const TileJob* d_jobs, const ListGroup& G, int* qnext)
{
	hipStream_t s = c->stream;
	constexpr int CX = T / 2 + 1, ROWS = 16, RG = 2, CPW = 4;
	constexpr int NCG = (CX + CPW - 1) / CPW;
	const TileJob* jobs = d_jobs + G.j0;
	if (G.max_nr > 0)
		PROF(c, "k_row_sums", k_row_sums_list<Tin><<<dim3(G.max_nr, G.nj), 256, 0, s>>>(jobs, c->rowsum.as<float>()));
	PROF(c, "k_block_rows", k_block_rows_list<<<dim3((G.max_np + 255) / 256, G.nj), 256, 0, s>>>(jobs,
		c->rowsum.as<float>(), c->bsums.as<float>()));
	// The row pass's items differ in cost (rows of 8-bit, 16-bit and float
	// micrographs, edge tiles), so it runs on a work queue; the column pass's
	// items are uniform and run one per workgroup (a queue measured slower).
	if (c->list_queue) {
		int gr = 0;
		CHK(resident_grid(c, k_tile_rows_queue<Tin, T, ROWS, RG>, (ROWS / 2) * (T / 8), G.items, &gr));
		PROF(c, "k_tile_rows_queue", k_tile_rows_queue<Tin, T, ROWS, RG><<<gr, (ROWS / 2) * (T / 8), 0, s>>>(jobs,
			G.nj, G.items, qnext + G.q, c->bsums.as<float>(), tt.mask.as<float>(), tt.tw.as<float2>(),
			c->rt.as<float2>()));
	} else {
		PROF(c, "k_tile_rows", k_tile_rows_list<Tin, T, ROWS, RG><<<G.items, (ROWS / 2) * (T / 8), 0, s>>>(jobs, G.nj,
			c->bsums.as<float>(), tt.mask.as<float>(), tt.tw.as<float2>(), c->rt.as<float2>()));
	}
	PROF(c, "k_tile_cols", k_tile_cols_list<T, CPW><<<dim3(NCG, G.max_split, G.nj), CPW * (T / 8),
		0, s>>>(jobs, c->rt.as<float2>(), tt.tw.as<float2>(), c->tpart.as<float>()));
	PROF(c, "k_tile_reduce", k_tile_reduce_list<<<dim3((T * CX + 255) / 256, G.nj), 256, 0, s>>>(jobs,
		c->tpart.as<float>(), T * CX, c->raw.as<float>()));
	HIPCHK(c, hipGetLastError());
	return GCTF_OK;
}

template <int T>
int launch_tile_list_tt(gctf_ctx* c, const TileTables& tt, const TileJob* d_jobs, const ListGroup& G, int* qnext)
{
	return G.f16 ? launch_tile_list_t<_Float16, T>(c, tt, d_jobs, G, qnext)
		: launch_tile_list_t<float, T>(c, tt, d_jobs, G, qnext);
}

int launch_tile_list(gctf_ctx* c, int T, const TileTables& tt, const TileJob* d_jobs, const ListGroup& G, int* qnext)
{
	switch (T) {
	case 256: return launch_tile_list_tt<256>(c, tt, d_jobs, G, qnext);
	case 512: return launch_tile_list_tt<512>(c, tt, d_jobs, G, qnext);
	case 1024: return launch_tile_list_tt<1024>(c, tt, d_jobs, G, qnext);
	default: return fail(c, GCTF_ERR_UNSUPPORTED, "tile size");
	}
}

// Per-micrograph plans of a list, copied from their uploaded table.
__global__ void k_copy_plans(ImgPlan* __restrict__ dst, const ImgPlan* __restrict__ src, int n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) dst[i] = src[i];
}

int gctf_estimate_list(gctf_ctx* ctx, const gctf_image* imgs, int n, gctf_result* out)
{
	if (!ctx) return GCTF_ERR_ARG;
	if (!imgs || !out || n < 1) return fail(ctx, GCTF_ERR_ARG, "null input/output or n < 1");
	HIPCHK(ctx, hipSetDevice(ctx->device));
	const int T = ctx->prm.tile_size;
	struct Item {
		RescalePlan rp;
		TileGrid g;
		SearchPlan* sp;
		int splits;
	};
	std::vector<Item> it((size_t)n);
	TileTables* tt;
	CHK(get_tile_tables(ctx, T, &tt));
	int kmax = 0;
	std::string key = "list";
	for (int i = 0; i < n; ++i) {
		const gctf_image& im = imgs[i];
		if (!im.data) return fail(ctx, GCTF_ERR_ARG, "null micrograph");
		if (im.dtype != GCTF_DTYPE_F32 && im.dtype != GCTF_DTYPE_F16) return fail(ctx, GCTF_ERR_ARG, "dtype");
		CHK(make_rescale_plan(ctx, im.nx, im.ny, im.pix, it[i].rp));
		const RescalePlan& rp = it[i].rp;
		if (rp.nnx < T || rp.nny < T) return fail(ctx, GCTF_ERR_ARG, "image smaller than one tile");
		it[i].g = make_grid(rp.nnx, rp.nny, T);
		if (it[i].g.ntx < 1 || it[i].g.nty < 1) return fail(ctx, GCTF_ERR_ARG, "no tile fits the image");
		CHK(get_search_plan(ctx, ctx->prm.kv, ctx->prm.cs_mm, ctx->prm.amp_contrast, rp.pixn, T, &it[i].sp));
		kmax = std::max(kmax, it[i].sp->ring.n);
		// tile splits: gctf_autotune's for this shape, else the context's
		auto tu = ctx->tuned.find(shape_key(im.nx, im.ny, im.dtype, im.pix));
		it[i].splits = tu == ctx->tuned.end() ? ctx->tile_splits : tu->second.splits;
		char k[128];
		snprintf(k, sizeof k, "/%p:%d:%d:%d:%a:%d", im.data, im.dtype, im.nx, im.ny, im.pix, it[i].splits);
		key += k;
	}
	CHK(ensure_image_buffers(ctx, n, T, kmax));
	ctx->kstride = kmax;
	// tile work list: micrographs read as float32 (rescaled ones and float32
	// inputs) first, then those read as float16; groups of consecutive jobs
	const int NRB = T / 16;
	std::vector<int> order;
	for (int pass = 0; pass < 2; ++pass)
		for (int i = 0; i < n; ++i) {
			const bool f16 = !it[i].rp.active && imgs[i].dtype == GCTF_DTYPE_F16;
			if ((int)f16 == pass) order.push_back(i);
		}
	std::vector<TileJob> jobs((size_t)n);
	std::vector<ListGroup> groups;
	std::vector<size_t> pad_off((size_t)n, 0), cbuf_off((size_t)n, 0);
	for (int k = 0; k < n; ++k) {
		const int i = order[k];
		const TileGrid& g = it[i].g;
		const RescalePlan& rp = it[i].rp;
		const int f16 = !rp.active && imgs[i].dtype == GCTF_DTYPE_F16;
		const int ntiles = g.ntx * g.nty;
		const size_t rt = (size_t)ntiles * (T / 2 + 1) * T;
		if (groups.empty() || groups.back().f16 != f16
			|| (groups.back().rt + rt) * sizeof(float2) > kListRtBytes) {
			ListGroup G;
			G.j0 = k; G.f16 = f16;
			G.q = (int)groups.size();
			groups.push_back(G);
		}
		ListGroup& G = groups.back();
		TileJob& J = jobs[(size_t)k];
		J.img = nullptr;
		J.pitch = rp.active ? rp.pitch : imgs[i].nx;
		J.g = g;
		J.rows_first = G.items;
		J.nsplit = std::min(it[i].splits, ntiles);
		J.need_rows = rp.active ? 0 : 1;
		J.raw_slot = i;
		const int nr = (g.nby - 1) * g.step + g.bsize;
		J.rowsum_off = (long long)G.rowsum; G.rowsum += (size_t)nr * g.nbx;
		J.sums_off = (long long)G.sums; G.sums += (size_t)g.nbx * g.nby * kSumParts;
		J.rt_off = (long long)G.rt; G.rt += rt;
		J.part_off = (long long)G.part; G.part += (size_t)J.nsplit * T * (T / 2 + 1);
		if (rp.active) {
			pad_off[(size_t)i] = G.pad; G.pad += (size_t)rp.pitch * rp.nny;
			cbuf_off[(size_t)i] = G.cbuf; G.cbuf += cbuf_elems(rp);
		}
		G.items += (ntiles * NRB + 1) / 2;
		if (J.need_rows) G.max_nr = std::max(G.max_nr, nr);
		G.max_np = std::max(G.max_np, g.nbx * g.nby * kSumParts);
		G.max_split = std::max(G.max_split, J.nsplit);
		G.nj += 1;
	}
	ListGroup most;
	for (const ListGroup& G : groups) {
		most.pad = std::max(most.pad, G.pad); most.rowsum = std::max(most.rowsum, G.rowsum);
		most.sums = std::max(most.sums, G.sums); most.rt = std::max(most.rt, G.rt);
		most.part = std::max(most.part, G.part); most.cbuf = std::max(most.cbuf, G.cbuf);
	}
	CHK(ensure_buf(ctx, ctx->cbufA, sizeof(float2) * std::max<size_t>(most.cbuf, 1)));
	CHK(ensure_buf(ctx, ctx->cbufB, sizeof(float2) * std::max<size_t>(most.cbuf, 1)));
	CHK(ensure_buf(ctx, ctx->padc, sizeof(float) * std::max<size_t>(most.pad, 1)));
	CHK(ensure_buf(ctx, ctx->rowsum, sizeof(float) * most.rowsum));
	CHK(ensure_buf(ctx, ctx->bsums, sizeof(float) * most.sums));
	CHK(ensure_buf(ctx, ctx->rt, sizeof(float2) * most.rt));
	CHK(ensure_buf(ctx, ctx->tpart, sizeof(float) * most.part));
	CHK(ensure_buf(ctx, ctx->jobs, sizeof(TileJob) * (size_t)n));
	CHK(ensure_buf(ctx, ctx->qnext, sizeof(int) * groups.size()));
	for (int k = 0; k < n; ++k) {
		const int i = order[k];
		jobs[(size_t)k].img = it[i].rp.active ? (const void*)(ctx->padc.as<float>() + pad_off[(size_t)i])
			: imgs[i].data;
	}
	// rescale jobs and their kernel classes, per group in stage order
	std::vector<RescaleJob> rjobs;
	std::vector<std::vector<RsClass>> gclasses(groups.size());
	std::vector<int> tab;
	for (size_t gi = 0; gi < groups.size(); ++gi) {
		const ListGroup& G = groups[gi];
		std::map<std::vector<int>, size_t> index;   // (stage, variant, maxn, f16) -> class
		auto& cls = gclasses[gi];
		for (int k = G.j0; k < G.j0 + G.nj; ++k) {
			const int i = order[k];
			const RescalePlan& rp = it[i].rp;
			if (!rp.active) continue;
			RescaleJob R;
			R.img = imgs[i].data;
			R.nx = rp.nx; R.ny = rp.ny; R.nnx = rp.nnx; R.nny = rp.nny; R.ncx = rp.nnx / 2 + 1;
			R.pitch = rp.pitch;
			R.rows = rp.rows->plan(); R.cf = rp.cf->plan(); R.ci = rp.ci->plan(); R.c2r = rp.c2r->plan();
			R.A = ctx->cbufA.as<float2>() + cbuf_off[(size_t)i];
			R.B = ctx->cbufB.as<float2>() + cbuf_off[(size_t)i];
			R.out = ctx->padc.as<float>() + pad_off[(size_t)i];
			R.rowsum = ctx->rowsum.as<float>() + jobs[(size_t)k].rowsum_off;
			R.g = it[i].g;
			const int rj = (int)rjobs.size();
			rjobs.push_back(R);
			int ck[4][3], items[4];
			CHK(rescale_classes(ctx, rp, imgs[i].dtype == GCTF_DTYPE_F16, ck, items));
			for (int st = 0; st < 4; ++st) {
				const std::vector<int> kk = {st, ck[st][0], ck[st][1], ck[st][2]};
				auto f = index.find(kk);
				if (f == index.end()) {
					RsClass C;
					C.stage = st; C.variant = ck[st][0]; C.maxn = ck[st][1]; C.f16 = ck[st][2];
					index[kk] = cls.size();
					cls.push_back(C);
					f = index.find(kk);
				}
				RsClass& C = cls[f->second];
				C.entries.push_back(rj);
				C.first.push_back(C.items);
				C.items += items[st];
			}
		}
		std::stable_sort(cls.begin(), cls.end(), [](const RsClass& a, const RsClass& b) { return a.stage < b.stage; });
		for (RsClass& C : cls) {
			C.first_off = tab.size();
			tab.insert(tab.end(), C.first.begin(), C.first.end());
			C.job_off = tab.size();
			tab.insert(tab.end(), C.entries.begin(), C.entries.end());
		}
	}
	// per-micrograph search plans, copied into place by one launch in the graph
	std::vector<ImgPlan> hplans((size_t)n);
	for (int i = 0; i < n; ++i) {
		const SearchPlan& sp = *it[i].sp;
		hplans[i].cc = sp.cc; hplans[i].ring = sp.ring; hplans[i].lo1d = sp.lo1d; hplans[i].hi1d = sp.hi1d;
	}
	CHK(ensure_buf(ctx, ctx->rjobs, sizeof(RescaleJob) * std::max<size_t>(rjobs.size(), 1)));
	CHK(ensure_buf(ctx, ctx->rtab, sizeof(int) * std::max<size_t>(tab.size(), 1)));
	CHK(ensure_buf(ctx, ctx->list_plans, sizeof(ImgPlan) * (size_t)n));
	// ordered on the context's stream (it may be a non-blocking one)
	HIPCHK(ctx, hipMemcpyAsync(ctx->jobs.p, jobs.data(), sizeof(TileJob) * (size_t)n, hipMemcpyHostToDevice,
		ctx->stream));
	if (!rjobs.empty())
		HIPCHK(ctx, hipMemcpyAsync(ctx->rjobs.p, rjobs.data(), sizeof(RescaleJob) * rjobs.size(), hipMemcpyHostToDevice,
			ctx->stream));
	if (!tab.empty())
		HIPCHK(ctx, hipMemcpyAsync(ctx->rtab.p, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice,
			ctx->stream));
	HIPCHK(ctx, hipMemcpyAsync(ctx->list_plans.p, hplans.data(), sizeof(ImgPlan) * (size_t)n, hipMemcpyHostToDevice,
		ctx->stream));
	HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
	char tail[64];
	snprintf(tail, sizeof tail, "/%d/%d/%d", ctx->tile_splits, ctx->cpb, ctx->list_queue);
	key += tail;
	auto record = [&]() -> int {
		const int nq = (int)groups.size();
		k_queue_reset<<<(nq + 63) / 64, 64, 0, ctx->stream>>>(ctx->qnext.as<int>(), nq);
		k_state_init<<<(n + 63) / 64, 64, 0, ctx->stream>>>(ctx->states.as<SearchState>(), n,
			ctx->prm.ext_phase_deg, std::max(ctx->prm.ext_phase_range_deg, 0.0f));
		for (size_t gi = 0; gi < groups.size(); ++gi) {
			for (const RsClass& C : gclasses[gi])
				CHK(launch_rs_class(ctx, C, ctx->rjobs.as<RescaleJob>(), ctx->rtab.as<int>()));
			CHK(launch_tile_list(ctx, T, *tt, ctx->jobs.as<TileJob>(), groups[gi], ctx->qnext.as<int>()));
		}
		CHK(launch_post(ctx, *tt, n));
		k_copy_plans<<<(n + 63) / 64, 64, 0, ctx->stream>>>(ctx->plans.as<ImgPlan>(), ctx->list_plans.as<ImgPlan>(), n);
		HIPCHK(ctx, hipGetLastError());
		CHK(launch_1d(ctx, *tt, 0, n, 0, 0.f, 0.f, ctx->prm.ext_phase_range_deg > 0.0f));
		CHK(launch_2d(ctx, T, 0, n, NX_BEGIN_DO2D, std::max(ctx->prm.ext_phase_range_deg, 0.0f)));
		return GCTF_OK;
	};
	CHK(run_graph(ctx, key, record));
	return read_results(ctx, n, out);
}

// ---- autotuner ---------------------------------------------------------------
int gctf_autotune(gctf_ctx* ctx, const void* dev_imgs, int dtype, int n, int nx, int ny, long long stride,
	float pix, int* best_chunk, int* best_splits, float* best_ms, int* best_search_block)
{
	if (!ctx) return GCTF_ERR_ARG;
	if (!dev_imgs || n < 1) return fail(ctx, GCTF_ERR_ARG, "null input or n < 1");
	if (stride < (long long)nx * ny) return fail(ctx, GCTF_ERR_ARG, "stride smaller than one image");
	HIPCHK(ctx, hipSetDevice(ctx->device));
	RescalePlan rp;
	CHK(make_rescale_plan(ctx, nx, ny, pix, rp));
	const int T = ctx->prm.tile_size;
	if (rp.nnx < T || rp.nny < T) return fail(ctx, GCTF_ERR_ARG, "image smaller than one tile");
	const TileGrid g = make_grid(rp.nnx, rp.nny, T);
	TileTables* tt;
	SearchPlan* sp;
	CHK(get_tile_tables(ctx, T, &tt));
	CHK(get_search_plan(ctx, ctx->prm.kv, ctx->prm.cs_mm, ctx->prm.amp_contrast, rp.pixn, T, &sp));
	CHK(ensure_image_buffers(ctx, n, T, sp->ring.n));
	const int chunks[] = {1, 2, 4, 8, 16};
	const int splits[] = {4, 8, 16};
	const int save_c = ctx->chunk, save_s = ctx->tile_splits;
	hipEvent_t a, b;
	HIPCHK(ctx, hipEventCreate(&a));
	HIPCHK(ctx, hipEventCreate(&b));
	float best = 1e30f;
	int bc = save_c, bs = save_s;
	int rc = GCTF_OK;
	for (int ci : chunks) {
		if (ci > n || rc != GCTF_OK) break;
		for (int si : splits) {
			ctx->chunk = ci;
			ctx->tile_splits = si;
			if ((rc = ensure_spectrum_buffers(ctx, g, T, ci)) != GCTF_OK) break;
			if (rp.active) {
				if ((rc = ensure_rescale_buffers(ctx, rp, ci)) != GCTF_OK) break;
				if ((rc = ensure_buf(ctx, ctx->padc, sizeof(float) * (size_t)rp.pitch * rp.nny * ci)) != GCTF_OK)
					break;
			}
			float tmin = 1e30f;
			for (int rep = 0; rep < 3 && rc == GCTF_OK; ++rep) {
				(void)hipEventRecord(a, ctx->stream);
				rc = record_spectra(ctx, rp, *tt, dev_imgs, dtype, n, stride);
				(void)hipEventRecord(b, ctx->stream);
				if (hipEventSynchronize(b) != hipSuccess) rc = fail(ctx, GCTF_ERR_HIP, "autotune timing");
				float ms = 0.f;
				(void)hipEventElapsedTime(&ms, a, b);
				if (rep > 0) tmin = std::min(tmin, ms);   // the first pass warms caches and plans
			}
			if (rc != GCTF_OK) break;
			if (tmin < best) { best = tmin; bc = ci; bs = si; }
		}
	}
	// search blocking: the plain 1D + 2D search of this batch's spectra with
	// 4, 8 and 16 candidates per evaluator workgroup
	const int save_b = ctx->cpb;
	int bb = save_b;
	if (rc == GCTF_OK) {
		ctx->chunk = bc;
		ctx->tile_splits = bs;
		rc = ensure_spectrum_buffers(ctx, g, T, std::min(bc, n));
		if (rc == GCTF_OK && rp.active) rc = ensure_rescale_buffers(ctx, rp, std::min(bc, n));
		if (rc == GCTF_OK && rp.active)
			rc = ensure_buf(ctx, ctx->padc, sizeof(float) * (size_t)rp.pitch * rp.nny * std::min(bc, n));
		ctx->kstride = sp->ring.n;
		if (rc == GCTF_OK) rc = record_spectra(ctx, rp, *tt, dev_imgs, dtype, n, stride);
		float sbest = 1e30f;
		const float p0 = ctx->prm.ext_phase_deg, prange = std::max(ctx->prm.ext_phase_range_deg, 0.0f);
		for (int cb : {4, 8, 16}) {
			if (rc != GCTF_OK) break;
			ctx->cpb = cb;
			float tmin = 1e30f;
			for (int rep = 0; rep < 3 && rc == GCTF_OK; ++rep) {
				(void)hipEventRecord(a, ctx->stream);
				k_state_init<<<(n + 63) / 64, 64, 0, ctx->stream>>>(ctx->states.as<SearchState>(), n, p0, prange);
				rc = fill_plans(ctx, 0, n, *sp);
				if (rc == GCTF_OK) rc = launch_1d(ctx, *tt, 0, n, 0, 0.f, 0.f, prange > 0.0f);
				if (rc == GCTF_OK) rc = launch_2d(ctx, T, 0, n, NX_BEGIN_DO2D, prange);
				(void)hipEventRecord(b, ctx->stream);
				if (hipEventSynchronize(b) != hipSuccess) rc = fail(ctx, GCTF_ERR_HIP, "autotune timing");
				float ms = 0.f;
				(void)hipEventElapsedTime(&ms, a, b);
				if (rep > 0) tmin = std::min(tmin, ms);
			}
			if (tmin < sbest) { sbest = tmin; bb = cb; }
		}
	}
	(void)hipEventDestroy(a);
	(void)hipEventDestroy(b);
	ctx->chunk = save_c;
	ctx->tile_splits = save_s;
	ctx->cpb = save_b;
	if (rc != GCTF_OK) return rc;
	ctx->tuned[shape_key(nx, ny, dtype, pix)] = gctf_ctx::Tuned{bc, bs, bb};
	ctx->tune_gen++;
	if (best_chunk) *best_chunk = bc;
	if (best_splits) *best_splits = bs;
	if (best_ms) *best_ms = best;
	if (best_search_block) *best_search_block = bb;
	return GCTF_OK;
}
