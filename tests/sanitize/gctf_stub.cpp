// gctf_stub.cpp -- CPU stand-in for the C-ABI entry points the GCtfFind
// command line and its facade call (include/gctf_hip.h), for builds of the
// host code under AddressSanitizer / UndefinedBehaviorSanitizer and
// ThreadSanitizer without a GPU.
//
// TEST INFRASTRUCTURE ONLY.  The arithmetic is a placeholder: every entry
// point reads and writes exactly the buffers the real one does (the caller's
// image rows, the padded image, the half and full spectra), so the
// sanitizers see the host code's buffer sizes, and the spectrum step takes a
// few milliseconds so the loader and writer threads run beside it.
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "gctf_hip.h"

struct gctf_ctx {
	gctf_params prm{};
	int nx = 0, ny = 0, nnx = 0, nny = 0, pitch = 0;
	float pix = 1.0f, pixn = 1.0f;
	std::vector<float> pad, half, raw;
	double mean = 0.0;
	float p0 = 0.0f, prange = 0.0f;
	gctf_result last{};
	std::string err;
	// gctf_rescale_prefetch: the image and its checksum when the copy began
	const float* pre_img = nullptr;
	size_t pre_n = 0;
	double pre_sum = 0.0;
};

extern "C" {

int gctf_create(const gctf_params* params, int device, gctf_ctx** out)
{
	if (!out) return GCTF_ERR_ARG;
	*out = nullptr;
	if (device != 0) return GCTF_ERR_HIP;
	gctf_ctx* c = new gctf_ctx;
	if (params) c->prm = *params;
	else c->prm.tile_size = 512;
	const size_t t = (size_t)c->prm.tile_size;
	c->half.assign((t / 2 + 1) * t, 0.0f);
	c->raw.assign((t / 2 + 1) * t, 0.0f);
	*out = c;
	return GCTF_OK;
}

void gctf_destroy(gctf_ctx* ctx) { delete ctx; }

const char* gctf_error_string(const gctf_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int gctf_synchronize(gctf_ctx* ctx) { return ctx ? GCTF_OK : GCTF_ERR_ARG; }

int gctf_host_alloc(size_t bytes, void** out)
{
	if (!out) return GCTF_ERR_ARG;
	*out = bytes ? std::malloc(bytes) : nullptr;
	return *out ? GCTF_OK : (bytes ? GCTF_ERR_NOMEM : GCTF_ERR_ARG);
}

void gctf_host_free(void* p) { std::free(p); }

int gctf_rescale_setup(gctf_ctx* ctx, int nx, int ny, float raw_pix, int* new_size, float* new_pix)
{
	if (!ctx || nx < 2 || ny < 2 || !(raw_pix > 0)) return GCTF_ERR_ARG;
	float bin = 1.2f / raw_pix;
	if (bin <= 1) bin = 1.0f;
	ctx->nx = nx; ctx->ny = ny; ctx->pix = raw_pix; ctx->pixn = raw_pix * bin;
	ctx->nnx = (int)((float)nx / bin + 0.5f) / 2 * 2;
	ctx->nny = (int)((float)ny / bin + 0.5f) / 2 * 2;
	ctx->pitch = (ctx->nnx / 2 + 1) * 2;
	ctx->pad.assign((size_t)ctx->pitch * ctx->nny, 0.0f);
	if (new_size) { new_size[0] = ctx->nnx; new_size[1] = ctx->nny; new_size[2] = ctx->pitch; }
	if (new_pix) *new_pix = ctx->pixn;
	return GCTF_OK;
}

// The real call copies the whole image asynchronously; the stand-in reads
// it all now, and the next gctf_rescale of the same image checks that it did
// not change in between (the caller's contract).
int gctf_rescale_prefetch(gctf_ctx* ctx, const float* img, int nx, int ny)
{
	if (!ctx || !img || nx < 1 || ny < 1) return GCTF_ERR_ARG;
	double s = 0.0;
	for (size_t i = 0; i < (size_t)nx * ny; ++i) s += img[i];
	ctx->pre_img = img;
	ctx->pre_n = (size_t)nx * ny;
	ctx->pre_sum = s;
	return GCTF_OK;
}

// Every pixel of the caller's nx x ny image is read; the padded image holds
// a nearest-neighbour copy of it.
int gctf_rescale(gctf_ctx* ctx, const float* img, int img_on_device)
{
	if (!ctx || !img || img_on_device) return GCTF_ERR_ARG;
	if (ctx->pre_img == img) {
		double c = 0.0;
		for (size_t i = 0; i < ctx->pre_n; ++i) c += img[i];
		ctx->pre_img = nullptr;
		if (c != ctx->pre_sum) {
			ctx->err = "prefetched image changed before its rescale";
			return GCTF_ERR_STATE;
		}
	}
	ctx->pre_img = nullptr;
	double s = 0.0;
	for (size_t i = 0; i < (size_t)ctx->nx * ctx->ny; ++i) s += img[i];
	ctx->mean = s / ((double)ctx->nx * ctx->ny);
	for (int y = 0; y < ctx->nny; ++y)
		for (int x = 0; x < ctx->nnx; ++x) {
			const int sx = (int)((long long)x * ctx->nx / ctx->nnx), sy = (int)((long long)y * ctx->ny / ctx->nny);
			ctx->pad[(size_t)y * ctx->pitch + x] = img[(size_t)sy * ctx->nx + sx];
		}
	return GCTF_OK;
}

int gctf_scaled_image(gctf_ctx* ctx, float** dev_pad_img)
{
	if (!ctx || !dev_pad_img) return GCTF_ERR_ARG;
	*dev_pad_img = ctx->pad.data();
	return GCTF_OK;
}

int gctf_setup1(gctf_ctx* ctx, float, float, float, float pix, float, float)
{
	if (!ctx || !(pix > 0)) return GCTF_ERR_ARG;
	ctx->pixn = pix;
	return GCTF_OK;
}

int gctf_setup2(gctf_ctx* ctx, int nx, int ny)
{
	if (!ctx || nx < ctx->prm.tile_size || ny < ctx->prm.tile_size) return GCTF_ERR_ARG;
	return GCTF_OK;
}

int gctf_set_phase(gctf_ctx* ctx, float p0, float range)
{
	if (!ctx) return GCTF_ERR_ARG;
	ctx->p0 = p0; ctx->prange = range;
	return GCTF_OK;
}

// Reads the whole padded image (rows of `pitch` floats) and fills the
// spectra; sleeps 3 ms in place of the GPU's work.
int gctf_gen_half_spectrum(gctf_ctx* ctx, const float* dev_pad_img)
{
	if (!ctx || !dev_pad_img) return GCTF_ERR_ARG;
	double s = 0.0;
	for (int y = 0; y < ctx->nny; ++y)
		for (int x = 0; x < ctx->nnx; ++x) s += std::fabs(dev_pad_img[(size_t)y * ctx->pitch + x]);
	const double m = s / ((double)ctx->nnx * ctx->nny);
	for (size_t i = 0; i < ctx->half.size(); ++i) {
		ctx->raw[i] = (float)(m * (1.0 + (double)(i % 257) / 257.0));
		ctx->half[i] = (float)std::sin(0.01 * (double)i) * (float)m;
	}
	std::this_thread::sleep_for(std::chrono::milliseconds(3));
	return GCTF_OK;
}

static int result(gctf_ctx* ctx, float dfm, gctf_result* out)
{
	gctf_result r;
	r.df_min = dfm - 100.0f; r.df_max = dfm + 100.0f;
	r.azimuth_deg = (float)std::fmod(ctx->mean * 1e3, 180.0) - 90.0f;
	r.ext_phase_deg = ctx->p0;
	r.score = 0.25f;
	r.ctf_res = 8.0f;
	ctx->last = r;
	if (out) *out = r;
	return GCTF_OK;
}

int gctf_do1d(gctf_ctx* ctx, gctf_result* out) { return ctx ? result(ctx, 15000.0f, out) : GCTF_ERR_ARG; }
int gctf_refine1d(gctf_ctx* ctx, float init_df, float, gctf_result* out)
{
	return ctx ? result(ctx, init_df, out) : GCTF_ERR_ARG;
}
int gctf_do2d(gctf_ctx* ctx, gctf_result* out)
{
	return ctx ? result(ctx, 15000.0f + 10.0f * (float)std::fabs(ctx->mean), out) : GCTF_ERR_ARG;
}
int gctf_refine(gctf_ctx* ctx, const float* df_mean, const float* ast_ratio, const float* ast_angle_deg,
	const float* ext_phase_deg, gctf_result* out)
{
	if (!ctx || !df_mean || !ast_ratio || !ast_angle_deg || !ext_phase_deg) return GCTF_ERR_ARG;
	return result(ctx, df_mean[0] + 0.01f * df_mean[1], out);
}

int gctf_spect_size(gctf_ctx* ctx, int* size, int half)
{
	if (!ctx || !size) return GCTF_ERR_ARG;
	const int t = ctx->prm.tile_size;
	size[0] = half ? t / 2 + 1 : t;
	size[1] = t;
	return GCTF_OK;
}

int gctf_get_half_spect(gctf_ctx* ctx, int raw, float* host_out)
{
	if (!ctx || !host_out) return GCTF_ERR_ARG;
	const std::vector<float>& v = raw ? ctx->raw : ctx->half;
	std::memcpy(host_out, v.data(), sizeof(float) * v.size());
	return GCTF_OK;
}

int gctf_set_half_spect(gctf_ctx* ctx, const float* host_spect)
{
	if (!ctx || !host_spect) return GCTF_ERR_ARG;
	std::memcpy(ctx->half.data(), host_spect, sizeof(float) * ctx->half.size());
	return GCTF_OK;
}

int gctf_full_spectrum(gctf_ctx* ctx, float* host_out)
{
	if (!ctx || !host_out) return GCTF_ERR_ARG;
	const int t = ctx->prm.tile_size;
	for (int i = 0; i < t * t; ++i) host_out[i] = ctx->half[(size_t)(i % (t / 2 + 1)) * t / 2];
	return GCTF_OK;
}

}  // extern "C"
