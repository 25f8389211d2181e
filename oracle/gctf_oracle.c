/*
 * gctf_oracle.c -- CPU restatement of GCtfFind's CTF-estimation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product library (libgctf_hip.so) never
 * links, loads or calls it.
 *
 * Arithmetic: every heavy numeric stage (FFTs, tile sums, box filters, CTF
 * values, correlation sums) runs in IEEE double.  Search control values
 * (candidate grids, step counts, ranges, stored CC scores) are IEEE single,
 * evaluated with the same operation order as the reference, because they
 * decide which candidate wins.  Build with -ffp-contract=off so no float
 * expression is fused.
 *
 * Each function cites the reference file:line it restates.  Quirks of the
 * reference that are reproduced or defined are listed next to the code.
 *
 * Pinning (what checks each function; the tests live in tests/):
 *   orc_fft              numpy.fft on random vectors (pow2, smooth, prime,
 *                        Bluestein sizes); inverse round trip
 *   orc_ctf_params       closed form lambda(300 kV) = 0.0196808 A
 *   orc_rescale          numpy rfft2 -> crop -> Re(ifft2) restated in numpy;
 *                        brute-force DFT on 12x10 -> 8x6; identity when b=1
 *   orc_tile_grid        hand-computed shape table (4096, 3412, 3838x3710 ...)
 *   orc_avg_spectrum     numpy rfft2 of the masked tile (independent code);
 *                        brute-force DFT on a 16x16 tile; Parseval
 *   orc_rm_background    brute-force box in numpy on a 16x... spectrum with
 *                        the wrap rules; constant-input invariant
 *   orc_lowpass          closed-form Gaussian transfer applied with numpy
 *   orc_radial_table     FP32-vs-FP64 difference positions; margin analysis
 *   orc_ctf1d/2d         closed-form CTF values, zeros at chi = k*pi
 *   orc_cc1d/2d          brute-force normalised dot in numpy; scale
 *                        invariance; CC = 1 when S equals the template
 *   search 1D/2D         recovery of known (df, astigmatism, azimuth) on
 *                        noiseless synthetic spectra; candidate replay
 *   orc_spectral_cc      brute-force Pearson per shell in numpy
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* FP64 FFT: iterative radix-2 for powers of two, Bluestein otherwise.       */
/* Unnormalised in both directions; forward uses exp(-2 pi i k n / N).       */
/* ------------------------------------------------------------------------ */
typedef struct { double re, im; } cplx;

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

static void fft_pow2(cplx* a, int n, int inverse)
{
	/* bit reversal */
	for (int i = 1, j = 0; i < n; i++) {
		int bit = n >> 1;
		for (; j & bit; bit >>= 1) j ^= bit;
		j ^= bit;
		if (i < j) { cplx t = a[i]; a[i] = a[j]; a[j] = t; }
	}
	double sgn = inverse ? 1.0 : -1.0;
	for (int len = 2; len <= n; len <<= 1) {
		int half = len >> 1;
		for (int k = 0; k < half; k++) {
			double ang = sgn * 2.0 * M_PI * (double)k / (double)len;
			double wr = cos(ang), wi = sin(ang);
			for (int i = k; i < n; i += len) {
				cplx u = a[i], v = a[i + half];
				double tr = v.re * wr - v.im * wi;
				double ti = v.re * wi + v.im * wr;
				a[i].re = u.re + tr; a[i].im = u.im + ti;
				a[i + half].re = u.re - tr; a[i + half].im = u.im - ti;
			}
		}
	}
}

static void fft_bluestein(cplx* a, int n, int inverse)
{
	int m = 1;
	while (m < 2 * n - 1) m <<= 1;
	cplx* w = (cplx*)malloc(sizeof(cplx) * n);
	cplx* x = (cplx*)calloc(m, sizeof(cplx));
	cplx* b = (cplx*)calloc(m, sizeof(cplx));
	double sgn = inverse ? 1.0 : -1.0;
	for (int k = 0; k < n; k++) {
		/* chirp angle pi k^2 / n, k^2 reduced mod 2n for accuracy */
		long long k2 = ((long long)k * k) % (2LL * n);
		double ang = sgn * M_PI * (double)k2 / (double)n;
		w[k].re = cos(ang); w[k].im = sin(ang);
	}
	for (int k = 0; k < n; k++) {
		x[k].re = a[k].re * w[k].re - a[k].im * w[k].im;
		x[k].im = a[k].re * w[k].im + a[k].im * w[k].re;
	}
	b[0].re = w[0].re; b[0].im = -w[0].im;
	for (int k = 1; k < n; k++) {
		b[k].re = w[k].re; b[k].im = -w[k].im;
		b[m - k] = b[k];
	}
	fft_pow2(x, m, 0);
	fft_pow2(b, m, 0);
	for (int k = 0; k < m; k++) {
		double re = x[k].re * b[k].re - x[k].im * b[k].im;
		double im = x[k].re * b[k].im + x[k].im * b[k].re;
		x[k].re = re; x[k].im = im;
	}
	fft_pow2(x, m, 1);
	double inv_m = 1.0 / (double)m;
	for (int k = 0; k < n; k++) {
		double re = x[k].re * inv_m, im = x[k].im * inv_m;
		a[k].re = re * w[k].re - im * w[k].im;
		a[k].im = re * w[k].im + im * w[k].re;
	}
	free(w); free(x); free(b);
}

static void fft_any(cplx* a, int n, int inverse)
{
	if (n <= 1) return;
	if (is_pow2(n)) fft_pow2(a, n, inverse);
	else fft_bluestein(a, n, inverse);
}

/* Exported for pinning against numpy: data = interleaved re,im. */
ORC_EXPORT void orc_fft(double* data, int n, int inverse)
{
	fft_any((cplx*)data, n, inverse);
}

/* ------------------------------------------------------------------------ */
/* a1: CTF parameters.  CCTFTheory.cpp:69-86, 283-287; GCalcCTF1D.cu:43-52;  */
/* GCalcCTF2D.cu:88-97.                                                      */
/* ------------------------------------------------------------------------ */
typedef struct {
	float wl;         /* wavelength in pixels          (CCTFTheory.cpp:77) */
	float cs;         /* Cs in pixels                  (CCTFTheory.cpp:78) */
	float ac;         /* amplitude contrast                                */
	float amp_phase;  /* atanf(ac/sqrtf(1-ac^2)), float (GCalcCTF2D.cu:94) */
	float ext_phase;  /* fmodf(phase, pi) radians      (CCTFTheory.cpp:85) */
	float pix;        /* pixel size A                                      */
} orc_ctf;

ORC_EXPORT float orc_wavelength(float kv)
{
	/* CCTFTheory.cpp:283-287: fKv*1000 is a float product, 0.9784*fKv
	 * promotes to double. */
	float kv1000 = kv * 1000.0f;
	double d = 12.26 / sqrt((double)kv1000 + 0.9784 * (double)kv * (double)kv);
	return (float)d;
}

static void ctf_setup(orc_ctf* p, float kv, float cs_mm, float ac, float pix,
	float ext_rad)
{
	float pi_f = (float)(4.0 * atan(1.0));
	p->wl = orc_wavelength(kv) / pix;
	p->cs = (float)((double)cs_mm * 1e7 / (double)pix);
	p->ac = ac;
	p->amp_phase = atanf(ac / sqrtf(1.0f - ac * ac));
	p->ext_phase = fmodf(ext_rad, pi_f);
	p->pix = pix;
}

/* out[0..4] = wavelength(pix), Cs(pix), amp phase (kernel form), ext phase,
 * wavelength(A). */
ORC_EXPORT void orc_ctf_params(float kv, float cs_mm, float ac, float pix,
	float ext_rad, float* out)
{
	orc_ctf p;
	ctf_setup(&p, kv, cs_mm, ac, pix, ext_rad);
	out[0] = p.wl; out[1] = p.cs; out[2] = p.amp_phase; out[3] = p.ext_phase;
	out[4] = orc_wavelength(kv);
}

/* ------------------------------------------------------------------------ */
/* a2: Fourier-crop rescale.  CRescaleImage.cpp:47-99, CCufft2D.cpp:23-114, */
/* GFtResize2D.cu:7-32,124-143.                                             */
/* ------------------------------------------------------------------------ */
ORC_EXPORT void orc_rescale_sizes(int nx, int ny, float pix, int* out_size,
	float* out_pix)
{
	float bin = 1.2f / pix;                      /* CRescaleImage.cpp:54 */
	if (bin <= 1) bin = 1.0f;
	float pixn = pix * bin;                      /* :56 (not rounding-corrected) */
	int nnx = (int)((float)nx / bin + 0.5f);     /* :58-61 */
	int nny = (int)((float)ny / bin + 0.5f);
	nnx = nnx / 2 * 2;
	nny = nny / 2 * 2;
	out_size[0] = nnx; out_size[1] = nny;
	out_size[2] = (nnx / 2 + 1) * 2;             /* padded pitch, :63 */
	out_pix[0] = pixn; out_pix[1] = bin;
}

/* img: nx*ny floats (contiguous rows).  out: nny rows of pitch
 * (nnx/2+1)*2 doubles; the two pad columns are left at zero.
 * When b > 1: F = unnormalised DFT2(img); keep kx < nnx/2+1, ky <= nny/2
 * from the top and the rest from the bottom (GFtResize2D.cu:15-31); the
 * output is Re(unnormalised inverse DFT2) of that half spectrum extended
 * by Hermitian symmetry (the C2R definition; its handling of the kx = 0 and
 * kx = nnx/2 columns is the library's and is not pinned by the reference). */
ORC_EXPORT void orc_rescale(const float* img, int nx, int ny, float pix,
	double* out)
{
	int sz[3]; float pp[2];
	orc_rescale_sizes(nx, ny, pix, sz, pp);
	int nnx = sz[0], nny = sz[1], pitch = sz[2];
	float bin = pp[1];
	memset(out, 0, sizeof(double) * (size_t)pitch * nny);
	if (bin == 1) {                              /* CRescaleImage.cpp:76-85 */
		for (int y = 0; y < nny; y++)
			for (int x = 0; x < nx; x++)
				out[(size_t)y * pitch + x] = img[(size_t)y * nx + x];
		return;
	}
	int cx = nx / 2 + 1, ncx = nnx / 2 + 1;
	/* forward: rows (real -> complex), keep ncx columns */
	cplx* spec = (cplx*)malloc(sizeof(cplx) * (size_t)ncx * ny);
	cplx* row = (cplx*)malloc(sizeof(cplx) * (size_t)(nx > ny ? nx : ny) * 2);
	for (int y = 0; y < ny; y++) {
		for (int x = 0; x < nx; x++) {
			row[x].re = img[(size_t)y * nx + x]; row[x].im = 0;
		}
		fft_any(row, nx, 0);
		for (int k = 0; k < ncx; k++) spec[(size_t)y * ncx + k] = row[k];
	}
	(void)cx;
	/* columns: forward length ny, crop to nny, inverse length nny */
	cplx* col = (cplx*)malloc(sizeof(cplx) * (size_t)(ny > nny ? ny : nny));
	cplx* g = (cplx*)malloc(sizeof(cplx) * (size_t)ncx * nny);
	for (int k = 0; k < ncx; k++) {
		for (int y = 0; y < ny; y++) col[y] = spec[(size_t)y * ncx + k];
		fft_any(col, ny, 0);
		cplx* cr = (cplx*)malloc(sizeof(cplx) * nny);
		for (int y = 0; y < nny; y++) {
			int ys = y;
			if (y > nny / 2) ys = y - nny + ny;  /* GFtResize2D.cu:20-22 */
			cr[y] = col[ys];
		}
		fft_any(cr, nny, 1);
		for (int y = 0; y < nny; y++) g[(size_t)y * ncx + k] = cr[y];
		free(cr);
	}
	/* rows: C2R of length nnx */
	cplx* z = (cplx*)malloc(sizeof(cplx) * nnx);
	for (int y = 0; y < nny; y++) {
		const cplx* gy = g + (size_t)y * ncx;
		for (int k = 0; k < ncx; k++) z[k] = gy[k];
		for (int k = ncx; k < nnx; k++) {
			z[k].re = gy[nnx - k].re; z[k].im = -gy[nnx - k].im;
		}
		fft_any(z, nnx, 1);
		for (int x = 0; x < nnx; x++) out[(size_t)y * pitch + x] = z[x].re;
	}
	free(z); free(g); free(col); free(row); free(spec);
}

/* ------------------------------------------------------------------------ */
/* a4: tile grid.  CGenAvgSpectrum.cpp:40-75, 123-139.  All integer.        */
/* out: [nTx, nTy, offX, offY, overlap, step]                               */
/* ------------------------------------------------------------------------ */
ORC_EXPORT void orc_tile_grid(int nx, int ny, int tile, int* out)
{
	int ov = (int)((float)tile * 0.50f);
	ov = ov / 2 * 2;
	int step = tile - ov;
	int ntx = (nx - ov) / step;
	int nty = (ny - ov) / step;
	out[0] = ntx; out[1] = nty;
	out[2] = (nx - ntx * tile + (ntx - 1) * ov) / 2;
	out[3] = (ny - nty * tile + (nty - 1) * ov) / 2;
	out[4] = ov; out[5] = step;
}

/* ------------------------------------------------------------------------ */
/* a7: round-edge mask value.  GRoundEdge.cu:14-24 with centre (T/2, T/2)   */
/* and size (T, T) (CGenAvgSpectrum.cpp:112-115).                           */
/* ------------------------------------------------------------------------ */
static double round_edge(int x, int y, int tile)
{
	double c = (double)((float)tile * 0.5f);
	double s = (double)((float)tile * 1.0f);
	double fx = 2.0 * fabs((double)x - c) / s;
	double fy = 2.0 * fabs((double)y - c) / s;
	double r = sqrt(fx * fx + fy * fy);
	if (r >= 1.0) return 0.0;
	double q = 0.5 * (1.0 - cos((double)3.1415926f * r));
	return 1.0 - q * q * q;
}

ORC_EXPORT void orc_round_edge_table(int tile, double* out)
{
	for (int y = 0; y < tile; y++)
		for (int x = 0; x < tile; x++) out[(size_t)y * tile + x] = round_edge(x, y, tile);
}

/* ------------------------------------------------------------------------ */
/* a5-a10: tile-averaged amplitude spectrum.                                 */
/* CGenAvgSpectrum.cpp:84-139 (tile loop, extraction), GCalcMoment2D.cu     */
/* (mean), GNormalize2D.cu:7-20 ((x-mean)/1), GRoundEdge.cu:7-25 (mask),     */
/* GCalcSpectrum.cu:15-31,90-105 (unnormalised R2C, |F|, DC to y = T/2),     */
/* GAddImages.cu:10-21 (avg = avg + tile * (1/n), tile order).              */
/* pad_img: rows of `pitch` doubles.  avg: (T/2+1) x T, row-major.          */
/* ------------------------------------------------------------------------ */
ORC_EXPORT void orc_avg_spectrum(const double* pad_img, int nx, int ny,
	int pitch, int tile, double* avg)
{
	int g[6];
	orc_tile_grid(nx, ny, tile, g);
	int ntx = g[0], nty = g[1], offx = g[2], offy = g[3], step = g[5];
	int ntiles = ntx * nty;
	int cmpx = tile / 2 + 1, cmpy = tile;
	double fac = (double)(1.0f / (float)ntiles);
	memset(avg, 0, sizeof(double) * (size_t)cmpx * cmpy);
	double* t = (double*)malloc(sizeof(double) * (size_t)tile * tile);
	double* mask = (double*)malloc(sizeof(double) * (size_t)tile * tile);
	orc_round_edge_table(tile, mask);
	cplx* buf = (cplx*)malloc(sizeof(cplx) * (size_t)cmpx * tile);
	cplx* line = (cplx*)malloc(sizeof(cplx) * tile);
	for (int it = 0; it < ntiles; it++) {
		int sx = offx + (it % ntx) * step;
		int sy = offy + (it / ntx) * step;
		double sum = 0.0;
		for (int y = 0; y < tile; y++)
			for (int x = 0; x < tile; x++) {
				double v = pad_img[(size_t)(sy + y) * pitch + sx + x];
				t[(size_t)y * tile + x] = v;
				sum += v;
			}
		double mean = sum / ((double)tile * tile);
		for (size_t i = 0; i < (size_t)tile * tile; i++) t[i] = (t[i] - mean) * mask[i];
		for (int y = 0; y < tile; y++) {
			for (int x = 0; x < tile; x++) { line[x].re = t[(size_t)y * tile + x]; line[x].im = 0; }
			fft_any(line, tile, 0);
			for (int k = 0; k < cmpx; k++) buf[(size_t)k * tile + y] = line[k];
		}
		for (int k = 0; k < cmpx; k++) {
			fft_any(buf + (size_t)k * tile, tile, 0);
			for (int y = 0; y < tile; y++) {
				cplx c = buf[(size_t)k * tile + y];
				int yd = (y + cmpy / 2) % cmpy;
				avg[(size_t)yd * cmpx + k] += sqrt(c.re * c.re + c.im * c.im) * fac;
			}
		}
	}
	free(line); free(buf); free(mask); free(t);
}

/* ------------------------------------------------------------------------ */
/* a11: 2D background removal.  GRmBackground2D.cu:11-94; called with       */
/* fMinFreq = 1/15 (CFindCtfBase.cpp:162-165).                              */
/* ------------------------------------------------------------------------ */
static int box_size(float min_freq, int cmpy)
{
	int b = (int)(min_freq * (float)cmpy);                 /* :33 */
	if ((float)b > 0.3f * (float)cmpy) b = (int)(0.3f * (float)cmpy);
	b = b / 2 * 2 + 1;
	if (b < 7) b = 7;
	return b;
}

/* Centre-zero test in float exactly as GRmBackground2D.cu:25-31:
 * fR = sqrtf((x*0.5/(cmpx-1))^2 + y^2/(cmpy^2)) compared with 0.04 (double). */
static int bg_is_centre(int x, int yc, int cmpx, int cmpy)
{
	float fr = (float)x * 0.5f / (float)(cmpx - 1);
	fr = sqrtf(fr * fr + (float)(yc * yc) / (float)(cmpy * cmpy));
	return (double)fr < 0.04;
}

ORC_EXPORT void orc_rm_background(const double* in, double* out, int log_spect,
	int cmpx, int cmpy, float min_freq)
{
	int half_x = cmpx - 1, half_y = cmpy / 2;
	int b = box_size(min_freq, cmpy), hb = b / 2;
	for (int y = 0; y < cmpy; y++) {
		int yc = y - half_y;
		for (int x = 0; x < cmpx; x++) {
			size_t i = (size_t)y * cmpx + x;
			if (bg_is_centre(x, yc, cmpx, cmpy)) { out[i] = 0.0; continue; }
			double acc = 0.0;
			for (int k = -hb; k <= hb; k++) {
				int yy = k + yc;
				for (int j = -hb; j <= hb; j++) {
					int xx = j + x, ix, iy;
					if (xx >= half_x) xx -= 2 * half_x;
					if (xx >= 0) { ix = xx; iy = half_y + yy; }
					else { ix = -xx; iy = half_y - yy; }
					if (iy < 0) iy += cmpy;
					else if (iy >= cmpy) iy -= cmpy;
					double v = in[(size_t)iy * cmpx + ix];
					acc += log_spect ? v : sqrt(v);
				}
			}
			acc /= (double)(b * b);
			if (!log_spect) acc *= acc;
			out[i] = in[i] - acc;
		}
	}
}

/* ------------------------------------------------------------------------ */
/* a12: flatness gate.  CFindCtfBase.cpp:167-174 via GCalcMoment2D.cu.      */
/* out: [mean, E[x^2], var]; the low-pass runs iff var >= 1.                */
/* ------------------------------------------------------------------------ */
ORC_EXPORT void orc_moments(const double* s, int n, double* out)
{
	double m1 = 0.0, m2 = 0.0;
	for (int i = 0; i < n; i++) { m1 += s[i]; m2 += s[i] * s[i]; }
	m1 /= n; m2 /= n;
	out[0] = m1; out[1] = m2; out[2] = m2 - m1 * m1;
}

/* The gate decision of CFindCtfBase.cpp:171-174: GCalcMoment2D returns the
 * mean and E[x^2] as floats, and fStd = fStd - fMean * fMean is evaluated in
 * IEEE single (compiled without contraction: the product is rounded before
 * the subtraction).  The moments themselves are the FP64 sums rounded once.
 * Returns 1 when the low-pass runs (fStd >= 1). */
ORC_EXPORT int orc_gate(const double* s, int n)
{
	double m[3];
	orc_moments(s, n, m);
	const float fmean = (float)m[0];
	float fstd = (float)m[1];
	fstd = fstd - fmean * fmean;
	return fstd < 1.0f ? 0 : 1;
}

/* ------------------------------------------------------------------------ */
/* a13: spectrum low-pass.  CFindCtfBase.cpp:191-219; GCalcSpectrum.cu:47-69 */
/* (unfold; row T read at yDst = 0 is taken mod T), GLowpass2D.cu:9-23,57-74 */
/* (exp(fScale*(kx^2+ky^2)), fScale = -10/(Nx^2+Ny^2)), forward normalised,  */
/* inverse unnormalised.  Output column x' = T/2 of the half spectrum is     */
/* periodic column 0 of the full plane (the reference reads the pad column   */
/* there; its value is not pinned).                                         */
/* ------------------------------------------------------------------------ */
ORC_EXPORT void orc_lowpass(double* half, int cmpx, int cmpy)
{
	int half_x = cmpx - 1, nxf = half_x * 2;
	cplx* f = (cplx*)malloc(sizeof(cplx) * (size_t)nxf * cmpy);
	for (int yd = 0; yd < cmpy; yd++)
		for (int xd = 0; xd < nxf; xd++) {
			int xs = xd - half_x, ys = yd;
			if (xs < 0) { xs = -xs; ys = (cmpy - yd) % cmpy; }
			f[(size_t)yd * nxf + xd].re = half[(size_t)ys * cmpx + xs];
			f[(size_t)yd * nxf + xd].im = 0.0;
		}
	cplx* line = (cplx*)malloc(sizeof(cplx) * (nxf > cmpy ? nxf : cmpy));
	for (int y = 0; y < cmpy; y++) fft_any(f + (size_t)y * nxf, nxf, 0);
	for (int x = 0; x < nxf; x++) {
		for (int y = 0; y < cmpy; y++) line[y] = f[(size_t)y * nxf + x];
		fft_any(line, cmpy, 0);
		for (int y = 0; y < cmpy; y++) f[(size_t)y * nxf + x] = line[y];
	}
	double dtemp = (double)nxf * nxf + (double)cmpy * cmpy;
	double scale = (double)(float)(-10.0f / dtemp);
	double norm = 1.0 / ((double)nxf * cmpy);
	for (int y = 0; y < cmpy; y++) {
		int ky = y > cmpy / 2 ? y - cmpy : y;
		for (int x = 0; x < nxf; x++) {
			int kx = x > nxf / 2 ? x - nxf : x;
			double h = exp(scale * (double)(kx * kx + ky * ky)) * norm;
			f[(size_t)y * nxf + x].re *= h;
			f[(size_t)y * nxf + x].im *= h;
		}
	}
	for (int x = 0; x < nxf; x++) {
		for (int y = 0; y < cmpy; y++) line[y] = f[(size_t)y * nxf + x];
		fft_any(line, cmpy, 1);
		for (int y = 0; y < cmpy; y++) f[(size_t)y * nxf + x] = line[y];
	}
	for (int y = 0; y < cmpy; y++) fft_any(f + (size_t)y * nxf, nxf, 1);
	for (int y = 0; y < cmpy; y++)
		for (int x = 0; x < cmpx; x++)
			half[(size_t)y * cmpx + x] = f[(size_t)y * nxf + (half_x + x) % nxf].re;
	free(line); free(f);
}

/* ------------------------------------------------------------------------ */
/* a14: radial average.  GRadialAvg.cu:11-33.  The sample table is built in  */
/* IEEE single with the reference's operation order; use_fma = 1 evaluates   */
/* the two multiply-adds fused (as a contracting device compiler would).     */
/* Row index cmpy (reached only at r = cmpx-1) is taken mod cmpy.           */
/* counts: cmpx ints; xy: 2 ints per sample (x, y), samples of r in order.   */
/* Returns the number of samples.                                           */
/* ------------------------------------------------------------------------ */
ORC_EXPORT int orc_radial_table(int cmpx, int cmpy, int use_fma, int* counts,
	int* xy)
{
	int n = 0;
	for (int r = 0; r < cmpx; r++) {
		float fp = use_fma ? fmaf(3.1415926f, (float)r, 0.5f)
		                   : 3.1415926f * (float)r + 0.5f;
		int np = (int)fp;
		if (np < 1) np = 1;
		else if (np > 180) np = 180;
		float step = 180.0f / (float)np;
		counts[r] = np;
		for (int i = 0; i < np; i++) {
			float a = use_fma ? fmaf((float)i, step, -90.0f)
			                  : -90.0f + (float)i * step;
			float rad = a * 0.017453f;
			float fx = use_fma ? fmaf((float)r, cosf(rad), 0.5f)
			                   : (float)r * cosf(rad) + 0.5f;
			float fy;
			if (use_fma) fy = fmaf((float)r, sinf(rad), (float)(cmpy / 2)) + 0.5f;
			else fy = (float)r * sinf(rad) + (float)(cmpy / 2) + 0.5f;
			int x = (int)fx, y = (int)fy;
			if (xy) { xy[2 * n] = x; xy[2 * n + 1] = y % cmpy; }
			n++;
		}
	}
	return n;
}

/* The same table evaluated in double (for the FP32-vs-FP64 comparison). */
ORC_EXPORT int orc_radial_table_f64(int cmpx, int cmpy, int* counts, int* xy,
	double* dist)
{
	int n = 0;
	for (int r = 0; r < cmpx; r++) {
		int np = (int)(3.1415926 * r + 0.5);
		if (np < 1) np = 1;
		else if (np > 180) np = 180;
		double step = 180.0 / np;
		counts[r] = np;
		for (int i = 0; i < np; i++) {
			double rad = (-90.0 + i * step) * 0.017453;
			double fx = r * cos(rad) + 0.5;
			double fy = r * sin(rad) + cmpy / 2 + 0.5;
			if (xy) { xy[2 * n] = (int)fx; xy[2 * n + 1] = ((int)fy) % cmpy; }
			if (dist) {
				double dx = fx - floor(fx), dy = fy - floor(fy);
				double d = dx < 1 - dx ? dx : 1 - dx;
				double e = dy < 1 - dy ? dy : 1 - dy;
				dist[n] = d < e ? d : e;
			}
			n++;
		}
	}
	return n;
}

ORC_EXPORT void orc_radial_avg(const double* half, int cmpx, int cmpy,
	double* avg)
{
	int* counts = (int*)malloc(sizeof(int) * cmpx);
	int nmax = orc_radial_table(cmpx, cmpy, 0, counts, NULL);
	int* xy = (int*)malloc(sizeof(int) * 2 * nmax);
	orc_radial_table(cmpx, cmpy, 0, counts, xy);
	int n = 0;
	for (int r = 0; r < cmpx; r++) {
		double s = 0.0;
		for (int i = 0; i < counts[r]; i++, n++)
			s += half[(size_t)xy[2 * n + 1] * cmpx + xy[2 * n]];
		avg[r] = s / counts[r];
	}
	free(xy); free(counts);
}

/* ------------------------------------------------------------------------ */
/* a15: CTF 1D.  GCalcCTF1D.cu:17-33,54-66.  df in pixels, add = amp phase  */
/* + extra phase (radians).                                                 */
/* ------------------------------------------------------------------------ */
ORC_EXPORT void orc_ctf1d(const float* prm, float df_pix, float add_phase,
	int n, double* out)
{
	double wl = prm[0], cs = prm[1];
	double pi1 = (double)3.141592654f;
	for (int i = 0; i < n; i++) {
		double s2 = ((double)i * 0.5) / ((double)n - 1.0);
		s2 *= s2;
		double chi = add_phase + pi1 * wl * s2 * (df_pix - 0.5 * wl * wl * s2 * cs);
		out[i] = -sin(chi);
	}
}

/* a16: CC 1D.  GCC1D.cu:15-58,96-125.  Bins x in [lo, hi) (Fourier pixels).*/
ORC_EXPORT float orc_cc1d(const double* ctf, const double* spec, int n,
	float lo, float hi, float bfac)
{
	double s = 0, s1 = 0, s2 = 0;
	for (int x = 0; x < n; x++) {
		if (!((float)x >= lo && (float)x < hi)) continue;
		double f = (double)x / ((double)n - 1.0);
		double c = (fabs(ctf[x]) - 0.5) * exp(-(double)bfac * f * f);
		s += c * spec[x]; s1 += c * c; s2 += spec[x] * spec[x];
	}
	if (s1 > 0 && s2 > 0) return (float)(s / sqrt(s1 * s2));
	return 0.0f;
}

/* ------------------------------------------------------------------------ */
/* a18: CTF 2D value at half-spectrum pixel (x, y).  GCalcCTF2D.cu:15-38.   */
/* ------------------------------------------------------------------------ */
static double ctf2d_at(int x, int y, int cmpx, int cmpy, double wl, double cs,
	double dfmean, double dfsigma, double az, double add_phase)
{
	double fx = (double)((float)x * 0.5f / (float)(cmpx - 1));
	double fy = (double)((float)(y - cmpy / 2) / (float)cmpy);
	double s2 = fx * fx + fy * fy;
	double th = atan(fy / (fx + (double)1e-30f));
	double df = dfmean + dfsigma * cos(2.0 * (th - az));
	double pi2 = (double)3.1415926f;
	return -sin(add_phase + pi2 * wl * s2 * (df - 0.5 * s2 * wl * wl * cs));
}

/* prm: [wl, cs]; dfmin/dfmax pixels; az radians. out: cmpx*cmpy. */
ORC_EXPORT void orc_ctf2d(const float* prm, float dfmin, float dfmax, float az,
	float add_phase, int cmpx, int cmpy, double* out)
{
	float dfmean = 0.5f * (dfmin + dfmax);       /* GCalcCTF2D.cu:104-105 */
	float dfsig = 0.5f * (dfmax - dfmin);
	for (int y = 0; y < cmpy; y++)
		for (int x = 0; x < cmpx; x++)
			out[(size_t)y * cmpx + x] = ctf2d_at(x, y, cmpx, cmpy, prm[0], prm[1],
				dfmean, dfsig, az, add_phase);
}

/* a19: ring membership, float exactly as GCC2D.cu:28-33,152-156. */
static void ring_bounds(float flo, float fhi, int cmpy, float* lo2, float* hi2)
{
	float l = flo / (float)cmpy, h = fhi / (float)cmpy;
	if (h > 0.75f) h = 0.75f;
	*lo2 = l * l; *hi2 = h * h;
}

static int in_ring(int x, int y, int cmpx, int cmpy, float lo2, float hi2,
	float* s2out)
{
	float fy = ((float)y - (float)cmpy * 0.5f) / (float)cmpy;
	float fx = (0.5f * (float)x) / (float)(cmpx - 1);
	fx = fx * fx + fy * fy;
	*s2out = fx;
	return !(fx < lo2 || fx > hi2);
}

/* Ring pixel list (for the bit-exact index contract test). */
ORC_EXPORT int orc_ring_pixels(int cmpx, int cmpy, float flo, float fhi,
	int* idx)
{
	float lo2, hi2, s2;
	ring_bounds(flo, fhi, cmpy, &lo2, &hi2);
	int n = 0;
	for (int y = 0; y < cmpy; y++)
		for (int x = 0; x < cmpx; x++)
			if (in_ring(x, y, cmpx, cmpy, lo2, hi2, &s2)) {
				if (idx) idx[n] = y * cmpx + x;
				n++;
			}
	return n;
}

/* CC 2D of a given CTF array and spectrum.  GCC2D.cu:13-61,71-99,144-170. */
ORC_EXPORT double orc_cc2d(const double* ctf, const double* spec, int cmpx,
	int cmpy, float flo, float fhi, float bfac)
{
	float lo2, hi2, s2f;
	ring_bounds(flo, fhi, cmpy, &lo2, &hi2);
	double s = 0, s1 = 0, s2 = 0;
	for (int y = 0; y < cmpy; y++)
		for (int x = 0; x < cmpx; x++) {
			if (!in_ring(x, y, cmpx, cmpy, lo2, hi2, &s2f)) continue;
			size_t i = (size_t)y * cmpx + x;
			double c = (fabs(ctf[i]) - 0.5) * exp(-(double)bfac * (double)s2f);
			s += c * spec[i]; s1 += c * c; s2 += spec[i] * spec[i];
		}
	double d = sqrt(s1 * s2);
	if (d == 0) return 0.0;
	return s / d;
}

/* ------------------------------------------------------------------------ */
/* a21: spectral CC shell.  GSpectralCC2D.cu:12-109,144-172.                 */
/* cc_out: cmpx values (may be NULL).  Returns the shell index.             */
/* ------------------------------------------------------------------------ */
ORC_EXPORT int orc_spectral_cc(const double* ctf, const double* spec, int cmpx,
	int cmpy, int width, double* cc_out)
{
	double* cc = (double*)malloc(sizeof(double) * cmpx);
	for (int k = 0; k < cmpx; k++) {
		int lo = k - width / 2, hi = lo + width;
		if (lo < 0) { lo = 0; hi = width; }
		else if (hi > cmpx) { lo = cmpx - width; hi = cmpx; }
		double sc = 0, ss = 0, scs = 0, scc = 0, sss = 0, cnt = 0;
		for (int y = 0; y < cmpy; y++) {
			float fy = fabsf((float)y - (float)cmpy * 0.5f);
			if (fy >= (float)hi) continue;
			for (int x = 0; x < cmpx; x++) {
				if (x >= hi) continue;
				float fr = sqrtf((float)(x * x) + fy * fy);
				if (fr < (float)lo || fr >= (float)hi) continue;
				size_t i = (size_t)y * cmpx + x;
				double c = ctf[i], s = spec[i];
				sc += c; ss += s; scs += c * s; scc += c * c; sss += s * s; cnt += 1;
			}
		}
		if (cnt == 0) { cc[k] = 0; continue; }
		sc /= cnt; ss /= cnt; scs /= cnt; scc /= cnt; sss /= cnt;
		scc -= sc * sc; sss -= ss * ss;
		if (scc <= 0 || sss <= 0) { cc[k] = 0; continue; }
		cc[k] = (scs - sc * ss) / sqrt(scc * sss);
	}
	int imax = 1;
	double fmaxv = -1e30;
	for (int i = 1; i < cmpx; i++) {
		if (cc[i] <= fmaxv) continue;
		fmaxv = cc[i]; imax = i;
	}
	int shell = imax;
	if (fmaxv >= 0.143) {
		for (int i = imax; i < cmpx; i++) {
			if (cc[i] < 0.143) break;
			shell = i;
		}
	}
	if (cc_out) memcpy(cc_out, cc, sizeof(double) * cmpx);
	free(cc);
	return shell;
}

/* ------------------------------------------------------------------------ */
/* The CFindCtf2D-shaped driver: per-micrograph state and the 1D/2D search. */
/* ------------------------------------------------------------------------ */
typedef struct {
	int stage;               /* see ORC_ST_* */
	float df_mean, ast, ang, phase;  /* A, ratio, deg, deg (1D: df, -, -, phase) */
	float cc;
} orc_trace;

enum { ORC_ST_1D = 0, ORC_ST_1D_REFINE = 1, ORC_ST_INIT = 2, ORC_ST_ASTIG = 3,
       ORC_ST_ASTMAG = 4, ORC_ST_ASTANG = 5, ORC_ST_DFMEAN = 6, ORC_ST_PHASE = 7 };

typedef struct {
	orc_ctf ctf;
	int tile, cmpx, cmpy, log_spect;
	float res[2];
	double* raw;       /* tile-averaged amplitude spectrum */
	double* spect;     /* background removed (+ low-passed) */
	double* radial;
	int lowpassed;
	/* CFindCtfBase results (CFindCTFInc.h:558-562) */
	float df_min, df_max, ast_ang, ext_phase, score, phase_range;
	/* CFindDefocus2D state (CFindCTFInc.h:512-529) */
	float d_dfmean, d_ast, d_ang, d_phase, d_ccmax, d_ctfres;
	float d_phrange[2], d_dfrange[2], d_astrange[2], d_angrange[2];
	float lo2d, hi2d;  /* GCC2D freq bounds (Fourier pixels) */
	/* ring cache for the 2D CC */
	int nring;
	int* ring_idx;
	double *ring_s2, *ring_th, *ring_w, *ring_spec;
	double ring_ss;
	/* 1D CC */
	float lo1d, hi1d;
	/* initial Do1D scan [lo, hi] A, step A (0: the reference scan) */
	float scan[3];
	/* candidate trace */
	orc_trace* trace;
	int ntrace, trace_cap, trace_on;
} orc_fc;

static void fc_trace(orc_fc* fc, int st, float a, float b, float c, float d, float cc)
{
	if (!fc->trace_on) return;
	if (fc->ntrace == fc->trace_cap) {
		fc->trace_cap = fc->trace_cap ? fc->trace_cap * 2 : 4096;
		fc->trace = (orc_trace*)realloc(fc->trace, sizeof(orc_trace) * fc->trace_cap);
	}
	orc_trace* t = &fc->trace[fc->ntrace++];
	t->stage = st; t->df_mean = a; t->ast = b; t->ang = c; t->phase = d; t->cc = cc;
}

static void fc_free_ring(orc_fc* fc)
{
	free(fc->ring_idx); free(fc->ring_s2); free(fc->ring_th); free(fc->ring_w);
	free(fc->ring_spec);
	fc->ring_idx = NULL; fc->ring_s2 = fc->ring_th = fc->ring_w = fc->ring_spec = NULL;
	fc->nring = 0;
}

ORC_EXPORT orc_fc* orc_fc_create(void)
{
	orc_fc* fc = (orc_fc*)calloc(1, sizeof(orc_fc));
	return fc;
}

ORC_EXPORT void orc_fc_destroy(orc_fc* fc)
{
	if (!fc) return;
	free(fc->raw); free(fc->spect); free(fc->radial); free(fc->trace);
	fc_free_ring(fc);
	free(fc);
}

/* CFindSeriesCtfs.cpp:116-131 (CCTFTheory::Setup with the rescaled pixel
 * size) + CFindCtfBase::Setup1 (CFindCtfBase.cpp:48-69) + CFindCtf1D/2D
 * Setup1.  The CFindDefocus2D extra phase, never initialised by the
 * reference before its first use (CFindDefocus2D.cpp:13-19,155), is 0. */
ORC_EXPORT void orc_fc_setup(orc_fc* fc, float kv, float cs, float ac,
	float pixn, int tile, int log_spect, float ext_phase_deg)
{
	float ext_rad = ext_phase_deg * 0.017453f;   /* CFindSeriesCtfs.cpp:122 */
	ctf_setup(&fc->ctf, kv, cs, ac, pixn, ext_rad);
	fc->tile = tile; fc->cmpy = tile; fc->cmpx = tile / 2 + 1;
	fc->log_spect = log_spect;
	float p = fc->ctf.pix;
	fc->res[0] = 25.0f * p;
	fc->res[1] = (2.0f * p) / 0.8f;
	if (fc->res[1] < 3.5f) fc->res[1] = 3.5f;
	size_t n = (size_t)fc->cmpx * fc->cmpy;
	free(fc->raw); free(fc->spect); free(fc->radial);
	fc->raw = (double*)calloc(n, sizeof(double));
	fc->spect = (double*)calloc(n, sizeof(double));
	fc->radial = (double*)calloc(fc->cmpx, sizeof(double));
	fc->df_min = fc->df_max = fc->ast_ang = fc->ext_phase = fc->score = 0.0f;
	fc->phase_range = 0.0f;
	fc->d_ast = 0.0f; fc->d_ang = 0.0f; fc->d_phase = 0.0f;
	fc_free_ring(fc);
	fc->ntrace = 0;
	/* CFindDefocus1D::mCorrelate (CFindDefocus1D.cpp:128-135) */
	float r1 = (float)((fc->cmpx - 1) * 2) * p;
	fc->lo1d = r1 / fc->res[0];
	fc->hi1d = r1 / fc->res[1];
	/* CFindDefocus2D::Setup2 (CFindDefocus2D.cpp:86-92) */
	float r2 = (float)fc->cmpy * p;
	fc->lo2d = r2 / fc->res[0];
	fc->hi2d = r2 / fc->res[1];
}

ORC_EXPORT void orc_fc_trace_enable(orc_fc* fc, int on) { fc->trace_on = on; fc->ntrace = 0; }
ORC_EXPORT int orc_fc_trace_count(orc_fc* fc) { return fc->ntrace; }
/* rows of 6 floats: stage, df_mean, ast, ang, phase, cc */
ORC_EXPORT void orc_fc_trace_get(orc_fc* fc, float* out)
{
	for (int i = 0; i < fc->ntrace; i++) {
		out[6 * i] = (float)fc->trace[i].stage;
		out[6 * i + 1] = fc->trace[i].df_mean;
		out[6 * i + 2] = fc->trace[i].ast;
		out[6 * i + 3] = fc->trace[i].ang;
		out[6 * i + 4] = fc->trace[i].phase;
		out[6 * i + 5] = fc->trace[i].cc;
	}
}

/* CFindCtfBase::mRemoveBackground + mLowpass (CFindCtfBase.cpp:158-219). */
static void fc_background(orc_fc* fc)
{
	size_t n = (size_t)fc->cmpx * fc->cmpy;
	orc_rm_background(fc->raw, fc->spect, fc->log_spect, fc->cmpx, fc->cmpy,
		1.0f / 15.0f);
	fc->lowpassed = 0;
	if (!orc_gate(fc->spect, (int)n)) return;    /* CFindCtfBase.cpp:174 */
	orc_lowpass(fc->spect, fc->cmpx, fc->cmpy);
	fc->lowpassed = 1;
}

/* CFindCtfBase::GenHalfSpectrum (CFindCtfBase.cpp:109-116) on a padded
 * rescaled image (rows of `pitch` doubles). */
ORC_EXPORT void orc_fc_gen_half_spectrum(orc_fc* fc, const double* pad_img,
	int nx, int ny, int pitch)
{
	orc_avg_spectrum(pad_img, nx, ny, pitch, fc->tile, fc->raw);
	fc_background(fc);
}

/* CFindCtfBase::SetHalfSpect: the background-removed spectrum directly. */
ORC_EXPORT void orc_fc_set_half_spect(orc_fc* fc, const double* spect)
{
	memcpy(fc->spect, spect, sizeof(double) * (size_t)fc->cmpx * fc->cmpy);
	fc_free_ring(fc);
}

ORC_EXPORT void orc_fc_set_raw_spect(orc_fc* fc, const double* raw)
{
	memcpy(fc->raw, raw, sizeof(double) * (size_t)fc->cmpx * fc->cmpy);
	fc_background(fc);
	fc_free_ring(fc);
}

ORC_EXPORT void orc_fc_get_spects(orc_fc* fc, double* raw, double* spect,
	double* radial)
{
	size_t n = (size_t)fc->cmpx * fc->cmpy;
	if (raw) memcpy(raw, fc->raw, sizeof(double) * n);
	if (spect) memcpy(spect, fc->spect, sizeof(double) * n);
	if (radial) memcpy(radial, fc->radial, sizeof(double) * fc->cmpx);
}

ORC_EXPORT int orc_fc_lowpassed(orc_fc* fc) { return fc->lowpassed; }

/* The initial Do1D scan: [lo, hi] A in steps of step A; all 0 restore
 * 3000 p^2 .. 30000 p^2 in 500 steps (CFindCtf1D.cpp:79-80). */
ORC_EXPORT void orc_fc_set_search1d(orc_fc* fc, float lo, float hi, float step)
{
	fc->scan[0] = lo; fc->scan[1] = hi; fc->scan[2] = step;
}

ORC_EXPORT void orc_fc_set_phase(orc_fc* fc, float p0, float range)
{
	fc->ext_phase = p0;                           /* CFindCtfBase.cpp:78-82 */
	fc->phase_range = range;
}

/* ---- 1D search: CFindDefocus1D.cpp:60-137 ------------------------------ */
static float fc_cc1d(orc_fc* fc, float df_a, float phase_deg)
{
	double ctf[4096];
	int n = fc->cmpx;
	float ext = phase_deg * 0.0174533f;              /* CFindDefocus1D.cpp:11,124 */
	float dfp = df_a / fc->ctf.pix;                  /* :125 */
	float add = fc->ctf.amp_phase + ext;             /* GCalcCTF1D.cu:63 */
	float prm[2] = { fc->ctf.wl, fc->ctf.cs };
	orc_ctf1d(prm, dfp, add, n, ctf);
	return orc_cc1d(ctf, fc->radial, n, fc->lo1d, fc->hi1d, 1.0f);
}

/* step > 0 replaces the reference step (range / 500, at least 10 A); the
 * point count keeps the reference rule (CFindDefocus1D.cpp:76-80). */
static void fc_search1d(orc_fc* fc, const float dfr[2], const float phr[2], float step,
	int stage, float* best_df, float* best_ph, float* best_cc)
{
	int ndf = 501;
	float dfrange = dfr[1] - dfr[0];
	float dfstep = dfrange / (float)(ndf - 1);
	if (dfstep < 10) dfstep = 10.0f;
	if (step > 0.0f) dfstep = step;
	ndf = (int)(dfrange / dfstep) / 2 * 2 + 1;
	int nps = 37;
	float psrange = phr[1] - phr[0];
	float psstep = psrange / (float)(nps - 1);
	if (psstep == 0) nps = 1;
	else nps = (int)(psrange / psstep) / 2 * 2 + 1;
	int npts = ndf * nps;
	float r0 = 0, r1 = 0, r2 = (float)-1e20;
	for (int i = 0; i < npts; i++) {
		int ifo = i % ndf, iph = i / ndf;
		float df = dfr[0] + (float)ifo * dfstep;
		float ph = phr[0] + (float)iph * psstep;
		if (ph < 0) ph = 0.0f;
		else if (ph > 150.0f) ph = 150.0f;
		float cc = fc_cc1d(fc, df, ph);
		fc_trace(fc, stage, df, 0, 0, ph, cc);
		if (cc > r2) { r0 = df; r1 = ph; r2 = cc; }
	}
	*best_df = r0; *best_ph = r1; *best_cc = r2;
}

/* CFindCtf1D::mCalcRadialAverage + mFindDefocus + mRefineDefocus
 * (CFindCtf1D.cpp:49-59,73-117). */
static void fc_find_defocus(orc_fc* fc)
{
	float p2 = fc->ctf.pix * fc->ctf.pix;
	float dfr[2] = { 3000.0f * p2, 30000.0f * p2 };   /* CFindCtf1D.cpp:79-80 */
	if (fc->scan[0] > 0.0f) { dfr[0] = fc->scan[0]; dfr[1] = fc->scan[1]; }
	float phr[2];
	phr[0] = fmaxf(fc->ext_phase - fc->phase_range / 2, 0.0f);
	phr[1] = fminf(phr[0] + fc->phase_range, 150.0f);
	float df, ph, cc;
	fc_search1d(fc, dfr, phr, fc->scan[2], ORC_ST_1D, &df, &ph, &cc);
	fc->ext_phase = ph; fc->df_min = df; fc->df_max = df; fc->score = cc;
}

static void fc_refine_defocus(orc_fc* fc, float range)
{
	float p = fc->ctf.pix;
	float mindf = 3000.0f * p * p;                  /* CFindCtf1D.cpp:97 */
	float dfr[2];
	dfr[0] = fmaxf(fc->df_min - range / 2, mindf);
	dfr[1] = dfr[0] + range;
	float prange = 0.2f * fc->phase_range;
	float phr[2] = { 0.0f, 0.0f };
	phr[0] = fmaxf(fc->ext_phase - prange / 2, 0.0f);
	phr[1] = fminf(phr[1] + prange, 150.0f);        /* quirk, CFindCtf1D.cpp:104 */
	float df, ph, cc;
	fc_search1d(fc, dfr, phr, 0.0f, ORC_ST_1D_REFINE, &df, &ph, &cc);
	fc->ext_phase = ph; fc->df_min = df; fc->df_max = df; fc->score = cc;
}

ORC_EXPORT void orc_fc_do1d(orc_fc* fc)
{
	orc_radial_avg(fc->spect, fc->cmpx, fc->cmpy, fc->radial);
	fc_find_defocus(fc);
	float range = fmaxf(0.3f * fc->df_min, 3000.0f);  /* CFindCtf1D.cpp:54 */
	fc_refine_defocus(fc, range);
}

ORC_EXPORT void orc_fc_refine1d(orc_fc* fc, float init_df, float range)
{
	fc->df_min = init_df; fc->df_max = init_df;       /* CFindCtf1D.cpp:61-72 */
	fc->score = (float)-1e20;
	orc_radial_avg(fc->spect, fc->cmpx, fc->cmpy, fc->radial);
	fc_refine_defocus(fc, range);
}

/* ---- 2D search: CFindDefocus2D.cpp:100-380 ----------------------------- */
static void fc_build_ring(orc_fc* fc)
{
	if (fc->nring) return;
	float lo2, hi2, s2f;
	ring_bounds(fc->lo2d, fc->hi2d, fc->cmpy, &lo2, &hi2);
	int n = orc_ring_pixels(fc->cmpx, fc->cmpy, fc->lo2d, fc->hi2d, NULL);
	fc->ring_idx = (int*)malloc(sizeof(int) * n);
	fc->ring_s2 = (double*)malloc(sizeof(double) * n);
	fc->ring_th = (double*)malloc(sizeof(double) * n);
	fc->ring_w = (double*)malloc(sizeof(double) * n);
	fc->ring_spec = (double*)malloc(sizeof(double) * n);
	int k = 0;
	fc->ring_ss = 0.0;
	for (int y = 0; y < fc->cmpy; y++)
		for (int x = 0; x < fc->cmpx; x++) {
			if (!in_ring(x, y, fc->cmpx, fc->cmpy, lo2, hi2, &s2f)) continue;
			double fx = (double)((float)x * 0.5f / (float)(fc->cmpx - 1));
			double fy = (double)((float)(y - fc->cmpy / 2) / (float)fc->cmpy);
			fc->ring_idx[k] = y * fc->cmpx + x;
			fc->ring_s2[k] = fx * fx + fy * fy;
			fc->ring_th[k] = atan(fy / (fx + (double)1e-30f));
			fc->ring_w[k] = exp(-16.0 * (double)s2f);   /* B = 16, :91 */
			fc->ring_spec[k] = fc->spect[fc->ring_idx[k]];
			fc->ring_ss += fc->ring_spec[k] * fc->ring_spec[k];
			k++;
		}
	fc->nring = n;
}

/* CFindDefocus2D::mCorrelate (CFindDefocus2D.cpp:363-380) with GCalcCTF2D
 * (GCalcCTF2D.cu:15-38,99-114) and GCC2D (GCC2D.cu). */
static float fc_correlate(orc_fc* fc, int stage, float ast, float ang, float ph)
{
	const float d2r = 0.01745329f;               /* CFindDefocus2D.cpp:11 */
	float ph_rad = ph * d2r;
	float az = ang * d2r;
	float dmin = fc->d_dfmean * (1.0f - ast) / fc->ctf.pix;   /* CFindCtfHelp.cpp:13-23 */
	float dmax = fc->d_dfmean * (1.0f + ast) / fc->ctf.pix;
	float mean = 0.5f * (dmin + dmax);
	float sig = 0.5f * (dmax - dmin);
	float add = fc->ctf.amp_phase + ph_rad;
	fc_build_ring(fc);
	double wl = fc->ctf.wl, cs = fc->ctf.cs, pi2 = (double)3.1415926f;
	double s = 0, s1 = 0;
	for (int k = 0; k < fc->nring; k++) {
		double s2 = fc->ring_s2[k];
		double df = (double)mean + (double)sig * cos(2.0 * (fc->ring_th[k] - (double)az));
		double ctf = -sin((double)add + pi2 * wl * s2 * (df - 0.5 * s2 * wl * wl * cs));
		double c = (fabs(ctf) - 0.5) * fc->ring_w[k];
		s += c * fc->ring_spec[k]; s1 += c * c;
	}
	double d = sqrt(s1 * fc->ring_ss);
	float cc = d == 0 ? 0.0f : (float)(s / d);
	fc_trace(fc, stage, fc->d_dfmean, ast, ang, ph, cc);
	return cc;
}

static void fc_find_astig(orc_fc* fc)
{
	const float tiny = (float)1e-20;
	float astr = fc->d_astrange[1] - fc->d_astrange[0];
	float angr = fc->d_angrange[1] - fc->d_angrange[0];
	if (astr < tiny && angr < tiny) return;
	int nast = 51, nang = 51;
	float aststep = astr / (float)nast;
	float angstep = angr / (float)nang;
	if ((double)aststep < 1e-5) nast = 1;
	if ((double)angstep < 1e-5) nang = 1;
	if (nast == 1 && nang == 1) return;
	float angmax = 0, astmax = 0, ccmax = (float)-1e20;
	for (int j = 0; j < nang; j++) {
		float ang = fc->d_angrange[0] + (float)j * angstep;
		for (int i = 0; i < nast; i++) {
			float ast = fc->d_astrange[0] + (float)i * aststep;
			float cc = fc_correlate(fc, ORC_ST_ASTIG, ast, ang, fc->d_phase);
			if (cc <= ccmax) continue;
			ccmax = cc; angmax = ang; astmax = ast;
		}
	}
	if (ccmax <= fc->d_ccmax) return;
	fc->d_ast = astmax; fc->d_ang = angmax; fc->d_ccmax = ccmax;
}

static void fc_refine_astmag(orc_fc* fc, float range)
{
	const float tiny = (float)1e-20;
	if (range < tiny) return;
	float mn = fmaxf(fc->d_ast - range * 0.5f, fc->d_astrange[0]);
	float mx = fminf(fc->d_ast + range * 0.5f, fc->d_astrange[1]);
	int n = 21;
	float step = (mx - mn) / (float)(n - 1);
	if (step < 0.001f) step = 0.001f;
	n = (int)((mx - mn) / step) / 2 * 2 + 1;
	if (n == 1) return;
	float best = 0, ccmax = (float)-1e20;
	for (int i = 0; i < n; i++) {
		float ast = mn + (float)i * step;
		float cc = fc_correlate(fc, ORC_ST_ASTMAG, ast, fc->d_ang, fc->d_phase);
		if (cc <= ccmax) continue;
		ccmax = cc; best = ast;
	}
	if (ccmax <= fc->d_ccmax) return;
	fc->d_ast = best; fc->d_ccmax = ccmax;
}

static void fc_refine_astang(orc_fc* fc, float range)
{
	const float tiny = (float)1e-20;
	if (range < tiny) return;
	float mn = fmaxf(fc->d_ang - range * 0.5f, fc->d_angrange[0]);
	float mx = fminf(fc->d_ang + range * 0.5f, fc->d_angrange[1]);
	int n = 51;
	float step = (mx - mn) / (float)(n - 1);
	if (step < 1.0f) step = 1.0f;
	n = (int)((mx - mn) / step) / 2 * 2 + 1;
	if (n == 1) return;
	float best = 0, ccmax = (float)-1e20;
	for (int i = 0; i < n; i++) {
		float ang = mn + (float)i * step;
		float cc = fc_correlate(fc, ORC_ST_ASTANG, fc->d_ast, ang, fc->d_phase);
		if (cc <= ccmax) continue;
		ccmax = cc; best = ang;
	}
	if (ccmax <= fc->d_ccmax) return;
	fc->d_ang = best; fc->d_ccmax = ccmax;
}

static void fc_refine_dfmean(orc_fc* fc, float range)
{
	const float tiny = (float)1e-30;
	if (range < tiny) return;
	float old = fc->d_dfmean;
	int n = 31;
	float mn = fmaxf(fc->d_dfmean - range * 0.5f, fc->d_dfrange[0]);
	float mx = fminf(fc->d_dfmean + range * 0.5f, fc->d_dfrange[1]);
	float step = (mx - mn) / (float)n;
	if (step < 20) step = 20.0f;
	n = (int)((mx - mn) / step) / 2 * 2 + 1;
	if (n == 1) return;
	float best = 0.0f, ccmax = (float)-1e20;
	for (int i = 0; i < n; i++) {
		fc->d_dfmean = mn + (float)i * step;
		float cc = fc_correlate(fc, ORC_ST_DFMEAN, fc->d_ast, fc->d_ang, fc->d_phase);
		if (cc <= ccmax) continue;
		best = fc->d_dfmean; ccmax = cc;
	}
	if (ccmax <= fc->d_ccmax) { fc->d_dfmean = old; return; }
	fc->d_dfmean = best; fc->d_ccmax = ccmax;
}

static void fc_refine_phase(orc_fc* fc, float range)
{
	const float tiny = (float)1e-30;
	if (range < tiny) return;
	int n = 61;
	float step = range / (float)n;
	if (step < 0.03f) return;
	if (step > 1.0f) {
		step = 1.0f;
		n = (int)(range / step);
		n = n / 2 * 2 + 1;
	}
	float mnp = fc->d_phrange[0] - fc->d_phrange[1] * 0.5f;
	float mxp = fc->d_phrange[0] + fc->d_phrange[1] * 0.5f;
	if (mnp < 0) mnp = 0.0f;
	if (mxp > 150) mxp = 150.0f;
	float ccmax = (float)-1e20, best = 0.0f;
	for (int i = 0; i < n; i++) {
		float ph = fc->d_phase + (float)(i - n / 2) * step;
		if (ph < mnp || ph > mxp) continue;
		float cc = fc_correlate(fc, ORC_ST_PHASE, fc->d_ast, fc->d_ang, ph);
		if (cc <= ccmax) continue;
		best = ph; ccmax = cc;
	}
	if (ccmax <= fc->d_ccmax) return;
	fc->d_phase = best; fc->d_ccmax = ccmax;
}

/* CFindDefocus2D::mIterate (CFindDefocus2D.cpp:153-183), including the
 * m_fCCMax = fOldCC reset after every astigmatism-magnitude step. */
static void fc_iterate(orc_fc* fc)
{
	fc->d_ccmax = fc_correlate(fc, ORC_ST_INIT, fc->d_ast, fc->d_ang, fc->d_phase);
	float old = fc->d_ccmax;
	fc_find_astig(fc);
	if (old > fc->d_ccmax) return;
	int iters = 20;
	fc->d_phase = fc->d_phrange[0];
	float dfr = fc->d_dfrange[1] - fc->d_dfrange[0];
	float astr = fc->d_astrange[1] - fc->d_astrange[0];
	float angr = fc->d_angrange[1] - fc->d_angrange[0];
	for (int i = 1; i < iters; i++) {
		float scale = 1.0f - (float)i * 0.5f / (float)iters;
		fc_refine_astmag(fc, scale * astr);
		if (old > fc->d_ccmax) return;
		fc->d_ccmax = old;
		fc_refine_astang(fc, scale * angr);
		fc_refine_dfmean(fc, scale * dfr);
		fc_refine_phase(fc, scale * fc->d_phrange[1]);
	}
}

/* CFindDefocus2D::mCalcCtfRes (CFindDefocus2D.cpp:382-400). */
static void fc_calc_ctf_res(orc_fc* fc)
{
	const float d2r = 0.01745329f;
	float ph_rad = fc->d_phase * d2r;
	float az = fc->d_ang * d2r;
	float dmin = fc->d_dfmean * (1.0f - fc->d_ast) / fc->ctf.pix;
	float dmax = fc->d_dfmean * (1.0f + fc->d_ast) / fc->ctf.pix;
	float prm[2] = { fc->ctf.wl, fc->ctf.cs };
	size_t n = (size_t)fc->cmpx * fc->cmpy;
	double* ctf = (double*)malloc(sizeof(double) * n);
	orc_ctf2d(prm, dmin, dmax, az, fc->ctf.amp_phase + ph_rad, fc->cmpx, fc->cmpy, ctf);
	int shell = orc_spectral_cc(ctf, fc->spect, fc->cmpx, fc->cmpy, 10, NULL);
	fc->d_ctfres = (float)fc->cmpy * fc->ctf.pix / (float)shell;
	free(ctf);
}

static void fc_get_results(orc_fc* fc)
{
	fc->df_min = fc->d_dfmean * (1.0f - fc->d_ast);   /* CFindDefocus2D.cpp:38-48 */
	fc->df_max = fc->d_dfmean * (1.0f + fc->d_ast);
	fc->ast_ang = fc->d_ang;
	fc->ext_phase = fc->d_phase;
	fc->score = fc->d_ccmax;
}

/* CFindCtf2D::Do2D (CFindCtf2D.cpp:44-53) -> CFindDefocus2D::DoIt
 * (CFindDefocus2D.cpp:110-124). */
ORC_EXPORT void orc_fc_do2d(orc_fc* fc)
{
	orc_fc_do1d(fc);
	float dfmean = (fc->df_min + fc->df_max) * 0.5f;
	fc->d_dfmean = dfmean; fc->d_ast = 0.0f; fc->d_ang = 0.0f;
	fc->d_phrange[0] = fc->ext_phase;
	fc->d_phrange[1] = fc->phase_range;
	fc->d_dfrange[0] = fc->d_dfmean * 0.9f;
	fc->d_dfrange[1] = fc->d_dfmean * 1.1f;
	fc->d_astrange[0] = 0.0f; fc->d_astrange[1] = 0.06f;
	fc->d_angrange[0] = -90.0f; fc->d_angrange[1] = 90.0f;
	fc_iterate(fc);
	fc_calc_ctf_res(fc);
	fc_get_results(fc);
}

/* CFindCtf2D::Refine (CFindCtf2D.cpp:55-68) -> CFindDefocus2D::Refine
 * (CFindDefocus2D.cpp:126-151), including the 0.02 astigmatism cap. */
ORC_EXPORT void orc_fc_refine(orc_fc* fc, const float* dfmean,
	const float* astratio, const float* astangle, const float* extphase)
{
	fc->d_dfmean = dfmean[0]; fc->d_ast = astratio[0]; fc->d_ang = astangle[0];
	fc->d_phrange[0] = extphase[0];
	float h = 0.5f * dfmean[1];
	fc->d_dfrange[0] = fmaxf(fc->d_dfmean - h, 3000.0f);
	fc->d_dfrange[1] = fc->d_dfmean + h;
	h = 0.5f * astratio[1];
	fc->d_astrange[0] = fmaxf(fc->d_ast - h, 0.0f);
	fc->d_astrange[1] = fminf(fc->d_ast + h, 0.02f);
	h = 0.5f * astangle[1];
	fc->d_angrange[0] = fmaxf(fc->d_ang - h, -90.0f);
	fc->d_angrange[1] = fminf(fc->d_ang + h, 90.0f);
	fc->d_phrange[1] = extphase[1];
	fc_iterate(fc);
	fc_calc_ctf_res(fc);
	fc_get_results(fc);
}

/* out: [df_min, df_max, azimuth deg, ext phase deg, score, ctf_res A] */
ORC_EXPORT void orc_fc_results(orc_fc* fc, float* out)
{
	out[0] = fc->df_min; out[1] = fc->df_max; out[2] = fc->ast_ang;
	out[3] = fc->ext_phase; out[4] = fc->score; out[5] = fc->d_ctfres;
}

/* ------------------------------------------------------------------------ */
/* f3: diagnostic full spectrum.  CFindCtfBase.cpp:118-133,                 */
/* CSpectrumImage.cpp:21-71, GCalcSpectrum.cu:47-69 (unpadded unfold),      */
/* GCalcCTF2D.cu:49-78,124-137 (abs CTF painted on the left half).          */
/* out: (cmpx-1)*2 x cmpy.                                                  */
/* ------------------------------------------------------------------------ */
ORC_EXPORT void orc_fc_full_spectrum(orc_fc* fc, double* out)
{
	int cmpx = fc->cmpx, cmpy = fc->cmpy, half_x = cmpx - 1, nxf = 2 * half_x;
	const double* h = fc->spect;
	for (int yd = 0; yd < cmpy; yd++)
		for (int xd = 0; xd < nxf; xd++) {
			int xs = xd - half_x, ys = yd;
			if (xs < 0) { xs = -xs; ys = (cmpy - yd) % cmpy; }
			out[(size_t)yd * nxf + xd] = h[(size_t)ys * cmpx + xs];
		}
	double m[3];
	orc_moments(h, cmpx * cmpy, m);
	double sd = m[2] < 0 ? 0.0 : sqrt(m[2]);
	/* CCTFTheory::SetDefocus / SetExtPhase (CCTFTheory.cpp:93-118) */
	float pi_f = (float)(4.0 * atan(1.0));
	float az = fc->ast_ang * 0.017453f;
	float dmin = fc->df_min / fc->ctf.pix, dmax = fc->df_max / fc->ctf.pix;
	float ext = fmodf(fc->ext_phase * 0.01745329f, pi_f);
	float prm[2] = { fc->ctf.wl, fc->ctf.cs };
	double* ctf = (double*)malloc(sizeof(double) * (size_t)cmpx * cmpy);
	orc_ctf2d(prm, dmin, dmax, az, fc->ctf.amp_phase + ext, cmpx, cmpy, ctf);
	float fmin = fc->ctf.pix / fc->res[0], fmax = fc->ctf.pix / fc->res[1];
	double gain = sd * 1.5;
	for (int y = 0; y < cmpy; y++)
		for (int bx = 0; bx < half_x; bx++) {
			float fy = ((float)y - (float)cmpy * 0.5f) / (float)cmpy;
			float fx = ((float)bx - (float)half_x) * 0.5f / (float)half_x;
			fx = sqrtf(fx * fx + fy * fy);
			if (fx < fmin || fx > fmax) continue;
			int ix = half_x - bx, iy = (cmpy - y) % cmpy;
			double v = fabs(ctf[(size_t)iy * cmpx + ix]);
			out[(size_t)y * nxf + bx] = (v - 0.5) * gain + m[0];
		}
	free(ctf);
}

/* ------------------------------------------------------------------------ */
/* Grid correlation (the exhaustive counterpart of the 2D search): FP64 CC  */
/* of every candidate m = (i_df * n_ast + i_ast) * n_az + i_az against one  */
/* spectrum, with the GCalcCTF2D / GCC2D formulas (GCalcCTF2D.cu:15-38,      */
/* GCC2D.cu:13-61) on the ring [1/res_lo, 1/res_hi] 1/A, B = 16.            */
/* Candidate values are IEEE single: df = df_lo + i*df_step etc., with      */
/* df_step = (df_hi - df_lo)/(n_df - 1), ast likewise, az_step =            */
/* (az_hi - az_lo)/n_az, and the CTF parameters as CFindDefocus2D::         */
/* mCorrelate forms them (CFindDefocus2D.cpp:363-380).                      */
/* ------------------------------------------------------------------------ */
ORC_EXPORT void orc_grid_cc(const double* spect, int cmpx, int cmpy, float kv, float cs, float ac,
	float pix, float ext_phase_deg, const float* gf, const int* gi, float res_lo, float res_hi,
	double* cc_out)
{
	orc_ctf p;
	ctf_setup(&p, kv, cs, ac, pix, 0.0f);
	const float df_lo = gf[0], df_hi = gf[1], ast_lo = gf[2], ast_hi = gf[3], az_lo = gf[4], az_hi = gf[5];
	const int n_df = gi[0], n_ast = gi[1], n_az = gi[2];
	const float df_step = n_df > 1 ? (df_hi - df_lo) / (float)(n_df - 1) : 0.0f;
	const float ast_step = n_ast > 1 ? (ast_hi - ast_lo) / (float)(n_ast - 1) : 0.0f;
	const float az_step = (az_hi - az_lo) / (float)n_az;
	float r2 = (float)cmpy * pix;
	float flo = r2 / res_lo, fhi = r2 / res_hi, lo2, hi2, s2f;
	ring_bounds(flo, fhi, cmpy, &lo2, &hi2);
	int nr = orc_ring_pixels(cmpx, cmpy, flo, fhi, NULL);
	double* rs2 = (double*)malloc(sizeof(double) * nr);
	double* rth = (double*)malloc(sizeof(double) * nr);
	double* rw = (double*)malloc(sizeof(double) * nr);
	double* rsp = (double*)malloc(sizeof(double) * nr);
	double ss = 0.0;
	int k = 0;
	for (int y = 0; y < cmpy; y++)
		for (int x = 0; x < cmpx; x++) {
			if (!in_ring(x, y, cmpx, cmpy, lo2, hi2, &s2f)) continue;
			double fx = (double)((float)x * 0.5f / (float)(cmpx - 1));
			double fy = (double)((float)(y - cmpy / 2) / (float)cmpy);
			rs2[k] = fx * fx + fy * fy;
			rth[k] = atan(fy / (fx + (double)1e-30f));
			rw[k] = exp(-16.0 * (double)s2f);
			rsp[k] = spect[(size_t)y * cmpx + x];
			ss += rsp[k] * rsp[k];
			k++;
		}
	const float d2r = 0.01745329f;
	const double wl = p.wl, csp = p.cs, pi2 = (double)3.1415926f;
	const int m = n_df * n_ast * n_az;
	for (int mm = 0; mm < m; mm++) {
		int iaz = mm % n_az, rest = mm / n_az, iast = rest % n_ast, idf = rest / n_ast;
		float dfm = df_lo + (float)idf * df_step;
		float ast = ast_lo + (float)iast * ast_step;
		float az = az_lo + (float)iaz * az_step;
		float ph_rad = ext_phase_deg * d2r, azr = az * d2r;
		float dmin = dfm * (1.0f - ast) / p.pix, dmax = dfm * (1.0f + ast) / p.pix;
		float mean = 0.5f * (dmin + dmax), sig = 0.5f * (dmax - dmin);
		float add = p.amp_phase + ph_rad;
		double s = 0.0, s1 = 0.0;
		for (int i = 0; i < nr; i++) {
			double df = (double)mean + (double)sig * cos(2.0 * (rth[i] - (double)azr));
			double c = -sin((double)add + pi2 * wl * rs2[i] * (df - 0.5 * rs2[i] * wl * wl * csp));
			c = (fabs(c) - 0.5) * rw[i];
			s += c * rsp[i]; s1 += c * c;
		}
		double d = sqrt(s1 * ss);
		cc_out[mm] = d == 0 ? 0.0 : s / d;
	}
	free(rs2); free(rth); free(rw); free(rsp);
}
