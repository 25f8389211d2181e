/*
 * gctf_hip.h -- C-ABI of libgctf_hip.so, the MI355X (gfx950) CTF-estimation
 * path of GCtfFind.
 *
 * Plain C types only: pointers, sizes, floats.  Every entry point returns a
 * GCTF_* status; gctf_error_string() describes the last failure of a
 * context.  One context drives one GPU; calls on one context must not
 * overlap (the context owns its device buffers and stream).
 *
 * Each declaration names the reference interface it replaces
 * (FindCTF/CFindCTFInc.h in the reference tree, file:line).  The
 * reference has no C-ABI; its call surface is the C++ classes
 * CRescaleImage, CCTFTheory, CFindCtfBase, CFindCtf1D and CFindCtf2D.
 * gctffind_amd/csrc/host/ re-creates those classes on top of this ABI.
 */
#ifndef GCTF_HIP_H
#define GCTF_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GCTF_API __attribute__((visibility("default")))

#define GCTF_OK 0
#define GCTF_ERR_ARG 1          /* invalid argument (null pointer, size, range) */
#define GCTF_ERR_HIP 2          /* a HIP runtime call failed */
#define GCTF_ERR_STATE 3        /* call order violated (e.g. Do2D before GenHalfSpectrum) */
#define GCTF_ERR_UNSUPPORTED 4  /* shape/tile outside what the kernels handle */
#define GCTF_ERR_NOMEM 5        /* device allocation failed */

#define GCTF_DTYPE_F32 0
#define GCTF_DTYPE_F16 1

typedef struct gctf_ctx gctf_ctx;

/* Command-line parameters of CInput (CInput.cpp:42-53, 113-179). */
typedef struct {
	float kv;                  /* -kV          (300)  */
	float cs_mm;               /* -Cs          (2.7)  */
	float amp_contrast;        /* -AmpContrast (0.07) */
	float ext_phase_deg;       /* -ExtPhase p0 (0)    */
	float ext_phase_range_deg; /* -ExtPhase range (0: no phase search) */
	int tile_size;             /* -TileSize    (512)  */
	int log_spect;             /* -LogSpect    (0)    */
} gctf_params;

/* CFindCtfBase public results (CFindCTFInc.h:558-562) plus
 * CFindDefocus2D::GetCtfRes (CFindCTFInc.h:499). */
typedef struct {
	float df_min;        /* A */
	float df_max;        /* A */
	float azimuth_deg;   /* astigmatism azimuth, dfMax axis from +x towards +y */
	float ext_phase_deg;
	float score;         /* CC */
	float ctf_res;       /* A (GSpectralCC2D shell) */
} gctf_result;

/* Creates a context on HIP device `device`; *out is null on failure. */
GCTF_API int gctf_create(const gctf_params* params, int device, gctf_ctx** out);
GCTF_API void gctf_destroy(gctf_ctx* ctx);
GCTF_API const char* gctf_error_string(const gctf_ctx* ctx);
/* Launch on a caller-owned hipStream_t (null: the context's own stream). */
GCTF_API int gctf_set_stream(gctf_ctx* ctx, void* hip_stream);
/* Wait for all work of the context. */
GCTF_API int gctf_synchronize(gctf_ctx* ctx);
/* Page-locked host memory (no reference counterpart) for callers that stage
 * micrographs on the host, e.g. the command line's loader threads: copies
 * from it to the device run at the full PCIe rate and without an extra
 * host copy.  GCTF_ERR_NOMEM when the allocation fails (*out null).  Free
 * with gctf_host_free once no copy from it is pending. */
GCTF_API int gctf_host_alloc(size_t bytes, void** out);
GCTF_API void gctf_host_free(void* p);

/* ---- CRescaleImage (CFindCTFInc.h:422-442) ---------------------------- */
/* CRescaleImage::Setup(int* piRawSize, float fRawPixSize) (:428).
 * new_size[0..2] = new nx, new ny, padded pitch (floats); *new_pix =
 * m_fPixSizeN. */
GCTF_API int gctf_rescale_setup(gctf_ctx* ctx, int nx, int ny, float raw_pix, int* new_size,
	float* new_pix);
/* CRescaleImage::DoIt(float* pfImage) (:429): nx*ny float32 image, host
 * memory when img_on_device == 0, else device memory. */
GCTF_API int gctf_rescale(gctf_ctx* ctx, const float* img, int img_on_device);
/* No reference counterpart: starts copying the nx*ny float32 host image img
 * to the device on a copy stream of the context and returns at once.  The
 * next gctf_rescale, if it is of the same pointer and size, uses that copy
 * instead of copying again; any other drops it.  For callers that estimate
 * one micrograph at a time (the command line): the next micrograph's PCIe
 * transfer then overlaps the current one's estimation.  img (page-locked
 * memory, e.g. from gctf_host_alloc, for an asynchronous copy) must stay
 * valid and unchanged until that gctf_rescale returns. */
GCTF_API int gctf_rescale_prefetch(gctf_ctx* ctx, const float* img, int nx, int ny);
/* CRescaleImage::GetScaledImg() (:430): device pointer, padded rows. */
GCTF_API int gctf_scaled_image(gctf_ctx* ctx, float** dev_pad_img);

/* ---- CCTFTheory + CFindCtf2D (CFindCTFInc.h:31-91, 541-615) ------------ */
/* CCTFTheory::Setup(kV, Cs, AmpContrast, PixelSize, AstTol, ExtPhase)
 * (:36-43) followed by CFindCtf2D::Setup1(CCTFTheory*) (:604). */
GCTF_API int gctf_setup1(gctf_ctx* ctx, float kv, float cs_mm, float amp_contrast, float pix,
	float ast_tol, float ext_phase_rad);
/* CFindCtfBase::Setup2(int* piImgSize) (:548). */
GCTF_API int gctf_setup2(gctf_ctx* ctx, int nx, int ny);
/* CFindCtfBase::SetPhase(float fInitPhase, float fPhaseRange) (:549). */
GCTF_API int gctf_set_phase(gctf_ctx* ctx, float p0_deg, float range_deg);

/* Initial defocus scan of Do1D: CFindCtf1D::mFindDefocus scans
 * 3000 p^2 .. 30000 p^2 A (p the working pixel size) in 500 steps
 * (CFindCtf1D.cpp:79-80, CFindDefocus1D.cpp:76-80); here [df_lo_A, df_hi_A]
 * in steps of step_A instead, with the reference's point-count rule
 * (int)(range / step) / 2 * 2 + 1.  All three 0 restore the reference scan.
 * Applies to the single-image and the batched plain searches. */
GCTF_API int gctf_set_search1d(gctf_ctx* ctx, float df_lo_A, float df_hi_A, float step_A);
/* CFindCtfBase::GenHalfSpectrum(float* gfPadImg) (:553): device image of
 * Setup2's size with row pitch (nx/2+1)*2 floats. */
GCTF_API int gctf_gen_half_spectrum(gctf_ctx* ctx, const float* dev_pad_img);
/* CFindCtf1D::Do1D() / Refine1D(float, float) (:588-589). */
GCTF_API int gctf_do1d(gctf_ctx* ctx, gctf_result* out);
GCTF_API int gctf_refine1d(gctf_ctx* ctx, float init_df, float df_range, gctf_result* out);
/* CFindCtf2D::Do2D() (:605). */
GCTF_API int gctf_do2d(gctf_ctx* ctx, gctf_result* out);
/* CFindCtf2D::Refine(afDfMean[2], afAstRatio[2], afAstAngle[2],
 * afExtPhase[2]) (:606-611); [0] = centre, [1] = full range. */
GCTF_API int gctf_refine(gctf_ctx* ctx, const float* df_mean, const float* ast_ratio,
	const float* ast_angle_deg, const float* ext_phase_deg, gctf_result* out);
/* No counterpart in the reference: gctf_do2d for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  CFindCtfBase::SetHalfSpect + Do2D
 * (:550, :605) -- Do1D, then the 2D search -- on n <= 16 background-removed
 * half spectra already on the device, d_spect [n][512][257] floats, at the
 * parameters of the last gctf_setup1 and the -ExtPhase values of
 * gctf_create; d_out gets n results in device memory.  Tile size 512 only.
 * Calls no hipMalloc/hipFree: its search buffers live in one static device
 * array per GPU (about 18 MB, held until the process ends), so calls that
 * overlap on one GPU, from any context, race.  Returns once the work is
 * enqueued on the context's stream (gctf_synchronize waits for it).
 * GCTF_ERR_ARG for n > 16; GCTF_ERR_STATE before gctf_setup1 or while
 * gctf_profile is on. */
GCTF_API int gctf_do2d_batch_static(gctf_ctx* ctx, const float* d_spect, int n, gctf_result* d_out);
/* CFindCtfBase::GetSpectSize(int*, bool bHalf) (:552). */
GCTF_API int gctf_spect_size(gctf_ctx* ctx, int* size, int half);
/* CFindCtfBase::GetHalfSpect(bool bRaw, bool bToHost=true) (:551): copies
 * (tile/2+1) x tile floats into host_out. */
GCTF_API int gctf_get_half_spect(gctf_ctx* ctx, int raw, float* host_out);
/* CFindCtfBase::SetHalfSpect(float*) (:550): host spectrum replaces the
 * background-removed spectrum. */
GCTF_API int gctf_set_half_spect(gctf_ctx* ctx, const float* host_spect);
/* CFindCtfBase::GenFullSpectrum() (:554): tile x tile floats, the fitted
 * |CTF| painted on the left half within the resolution ring. */
GCTF_API int gctf_full_spectrum(gctf_ctx* ctx, float* host_out);

/* ---- index tables ----------------------------------------------------
 * The integer contracts the spectrum and search stages are built on, as
 * the library computes them.  The reference forms them inside its kernels
 * and setup code; these queries expose the library's own tables. */
/* CGenAvgSpectrum::SetSizes (CGenAvgSpectrum.cpp:40-75) for an nx x ny
 * image and tile size `tile`: out6 = tiles in x, tiles in y, x offset,
 * y offset, overlap, step (tile i at offset + (i % ntx, i / ntx) * step). */
GCTF_API int gctf_tile_grid(int nx, int ny, int tile, int* out6);
/* Pixels y * (tile/2+1) + x of the half spectrum in the 2D correlation
 * ring (GCC2D.cu:28-33, 152-156) at working pixel size `pix` A: the
 * default ring [25 pix, max(2 pix / 0.8, 3.5)] A (CFindCtfBase.cpp:64-68)
 * when res_lo = res_hi = 0, else [res_lo, res_hi] A.  *n gets the count;
 * the first min(max, count) indices go to idx (host), in the order the
 * search reads them. */
GCTF_API int gctf_ring_pixels(gctf_ctx* ctx, float pix, float res_lo, float res_hi, int* idx, int max, int* n);
/* GRadialAvg::DoIt (GRadialAvg.cu:11-53) on n device half spectra
 * [n][tile][tile/2+1]: d_radial [n][ld] device floats, bins 0 .. tile/2,
 * ld >= tile/2 + 1.  Returns after the averages are written. */
GCTF_API int gctf_radial_average(gctf_ctx* ctx, const float* d_spect, int n, float* d_radial, int ld);

/* ---- Batched path (no reference counterpart) --------------------------
 * Rescale + GenHalfSpectrum + Do2D (the reference's full-search package,
 * CFindSeriesCtfs.cpp:108-141) for n micrographs of one shape and pixel
 * size, device-resident, image i at dev_imgs + i * stride elements of
 * `dtype`.  The whole pipeline of a batch is captured once into a HIP graph
 * and replayed.  out: n results (host memory). */
GCTF_API int gctf_estimate_batch(gctf_ctx* ctx, const void* dev_imgs, int dtype, int n, int nx,
	int ny, long long stride, float pix, gctf_result* out);

/* Exhaustive grid search on MFMA (no reference counterpart; it replaces
 * the coordinate descent of CFindDefocus2D::mIterate with a full grid).
 * Candidate m = (i_df * n_ast + i_ast) * n_az + i_az with mean defocus
 * df_lo + i_df (df_hi - df_lo)/(n_df - 1) A, astigmatism ratio
 * ast_lo + i_ast (ast_hi - ast_lo)/(n_ast - 1), azimuth
 * az_lo + i_az (az_hi - az_lo)/n_az degrees.  The correlation ring is
 * [1/res_lo, 1/res_hi] 1/A with the GCC2D weighting (CFindDefocus2D.cpp:91). */
typedef struct {
	float df_lo, df_hi;
	int n_df;
	float ast_lo, ast_hi;
	int n_ast;
	float az_lo, az_hi;
	int n_az;
	float res_lo, res_hi;     /* A, res_lo > res_hi */
} gctf_grid;

/* Spectrum stages as gctf_estimate_batch, then CC of every grid candidate
 * against every micrograph as a BF16 GEMM with FP32 accumulation, the best
 * 16 per micrograph re-scored in FP32 (FP64 sums); out gets the best
 * re-scored candidate.  n <= 256.  Optional host outputs: cc_out [n][M]
 * BF16-GEMM CCs, top_out [n][16] candidate indices, rescored_out [n][16]. */
GCTF_API int gctf_grid_search_batch(gctf_ctx* ctx, const void* dev_imgs, int dtype, int n, int nx,
	int ny, long long stride, float pix, const gctf_grid* grid, gctf_result* out, float* cc_out,
	int* top_out, float* rescored_out);

/* Knobs of the batched path (0 keeps the current value): images per
 * spectrum chunk, tile-column splits, graph use (1 on, -1 off), 2D search
 * blocking (candidates per evaluator workgroup: 2, 4, 8 or 16) and the tile
 * row pass of gctf_estimate_list (1: persistent workgroups on a work queue,
 * -1: one workgroup per item).  None of them changes results except the
 * tile splits (FP32 partial-sum grouping of the tile average). */
GCTF_API int gctf_tune(gctf_ctx* ctx, int chunk, int tile_splits, int use_graph, int search_block,
	int list_queue);

/* Kernel timing (no reference counterpart): while enabled, every kernel the
 * context launches is bracketed by HIP events on its stream and batches run
 * without graphs.  gctf_profile(ctx, 1) starts and clears; the report
 * aggregates per kernel name since the start: names (64 chars each),
 * total milliseconds and launch counts, at most `max` rows. */
GCTF_API int gctf_profile(gctf_ctx* ctx, int enable);
GCTF_API int gctf_profile_report(gctf_ctx* ctx, int max, char* names, double* total_ms,
	int* count, int* nout);

/* ---- batching (no reference counterpart) ---------------------------------- */

/* One micrograph of an uneven batch. */
typedef struct {
	const void* data;         /* device address, row-major, nx fastest */
	int dtype;                /* GCTF_DTYPE_F32 or GCTF_DTYPE_F16 */
	int nx, ny;
	float pix;                /* A */
} gctf_image;

/* Micrographs of different shapes, types and pixel sizes in one call (one
 * HIP graph per list): each is rescaled and tile-averaged on its own plan,
 * then background removal, low-pass and the plain 1D + 2D search run over
 * the whole list with per-micrograph CTF constants and rings.  Results
 * equal gctf_estimate_batch's on each micrograph. */
GCTF_API int gctf_estimate_list(gctf_ctx* ctx, const gctf_image* imgs, int n, gctf_result* out);

/* A tilt series of n same-shape micrographs in one call (CFindSeriesCtfs,
 * CFindSeriesCtfs.cpp:133-203): spectra of all images batched, then the
 * zero-tilt image (CInputFolder::FindZeroTilt, tomo != 0; image 0 otherwise)
 * gets the full search and every other image, in index order, a Refine
 * around the best-scoring image so far -- defocus range max(0.5 mean,
 * 5000 pix^2), astigmatism and azimuth fixed when tomo -- with a full search
 * when its score is <= 0.  The chain runs on the device, in one graph. */
GCTF_API int gctf_estimate_series(gctf_ctx* ctx, const void* dev_imgs, int dtype, int n, int nx, int ny,
	long long stride, float pix, const float* tilts_deg, int tomo, gctf_result* out);

/* Time the spectrum stages of this batch for every (images per chunk, tile
 * splits) pair in {1, 2, 4, 8, 16} x {4, 8, 16}, then the plain search of
 * its spectra with 4, 8 and 16 candidates per 2D-evaluator workgroup, and
 * keep the fastest for this shape (nx, ny, dtype, pix): later batch,
 * series, learned and list calls of that shape use them (results do not
 * depend on the search blocking).  Outputs optional; best_ms is the
 * spectrum stages' time. */
GCTF_API int gctf_autotune(gctf_ctx* ctx, const void* dev_imgs, int dtype, int n, int nx, int ny,
	long long stride, float pix, int* best_chunk, int* best_splits, float* best_ms, int* best_search_block);

/* ---- learned-proposal path (no reference counterpart) -------------------
 * A 1D convolutional network reads each micrograph's radial spectrum (the
 * CFindCtf1D input, CFindCtf1D.cpp:113-117) and proposes its mean defocus
 * with a confidence.  Confident proposals seed a narrowed 2D search (the
 * CFindCtf2D::Refine hook, CFindCtf2D.cpp:55-68, with DoIt's astigmatism
 * box, CFindDefocus2D.cpp:113-118, instead of Refine's 0.02 cap); the
 * others, and narrowed searches that fail or end at the window edge, get
 * the full plain search (the reference's score <= 0 fallback,
 * CFindSeriesCtfs.cpp:144). */

/* Candidate counts of a 2D search program.  The reference's (the plain
 * path): 51 x 51 grid, 21 / 51 / 31 line candidates, 20 rounds, 20 A floor,
 * no fine stage. */
typedef struct {
	int n_ast, n_ang;     /* astigmatism x azimuth grid of the first stage */
	int n_mag, n_angl;    /* astigmatism-magnitude / azimuth line candidates */
	int n_df;             /* mean-defocus line candidates */
	int iters;            /* rounds (mIterate's iIterations) */
	float df_minstep;     /* mean-defocus step floor, A */
	int fine_n;           /* final mean-defocus stage: candidates (0: none) */
	float fine_step;      /* its step, A */
} gctf_search_program;

typedef struct {
	float df_mean;        /* proposed mean defocus, A */
	float astig;          /* proposed df_max - df_min, A */
	float azimuth_deg;    /* proposed azimuth (a radial spectrum carries none; informative only) */
	float confidence;     /* in [0, 1] */
	int fallback;         /* 1: the plain search produced this micrograph's result */
} gctf_proposal;

/* Defaults of the learned path's narrowed search program. */
GCTF_API int gctf_learned_program_default(gctf_search_program* out);

/* Spectrum stages as gctf_estimate_batch, then the radial average of every
 * micrograph (CFindCtf1D.cpp:113-117): d_radial_out [n][ld] device floats,
 * bins 0 .. min(ld, tile/2 + 1) - 1. */
GCTF_API int gctf_radial_batch(gctf_ctx* ctx, const void* dev_imgs, int dtype, int n, int nx, int ny,
	long long stride, float pix, float* d_radial_out, int ld);

/* Install the proposal network's parameters (gctf_net_param_count() FP32
 * values, host memory, the layout gctf_net_get_params returns). */
GCTF_API int gctf_set_proposal_params(gctf_ctx* ctx, const float* host_params, long long count);

/* The learned path for a batch of device-resident micrographs: spectrum
 * stages, proposal network (BF16 MFMA), the narrowed search of `prog`
 * (NULL: gctf_learned_program_default) in [df - window_A, df + window_A],
 * fallback to the plain search below conf_threshold.  prop_out optional. */
GCTF_API int gctf_estimate_batch_learned(gctf_ctx* ctx, const void* dev_imgs, int dtype, int n, int nx,
	int ny, long long stride, float pix, const gctf_search_program* prog, float conf_threshold,
	float window_A, gctf_result* out, gctf_proposal* prop_out);

/* ---- proposal-network trainer (no reference counterpart) ----------------
 * Forward and backward passes on MFMA (bf16 = 1: BF16 operands, FP32
 * accumulation and master parameters; bf16 = 0: FP32 operands, for gradient
 * checks), AdamW on the FP32 parameters.  Device pointers are read on the
 * trainer's own stream: synchronise the producer first. */
typedef struct gctf_net gctf_net;
GCTF_API long long gctf_net_param_count(void);
GCTF_API int gctf_net_create(int device, int max_batch, gctf_net** out);
GCTF_API void gctf_net_destroy(gctf_net* net);
GCTF_API const char* gctf_net_error_string(const gctf_net* net);
GCTF_API int gctf_net_set_params(gctf_net* net, const float* host, long long count);
GCTF_API int gctf_net_get_params(gctf_net* net, float* host, long long count);
GCTF_API int gctf_net_reset_optimizer(gctf_net* net);
/* raw outputs [n][4]: ln(mean df / 1e4 A), astigmatism / 1e3 A, azimuth rad, confidence logit */
GCTF_API int gctf_net_forward(gctf_net* net, const float* d_radial, int ld, int n, int bf16, float* host_out);
/* d_target [n][4] device: mean df A, astigmatism A, azimuth deg, unused.
 * Leaves the gradient for gctf_net_adam_step; loss / host_grad optional. */
GCTF_API int gctf_net_loss_grad(gctf_net* net, const float* d_radial, int ld, const float* d_target, int n,
	int bf16, double* loss, float* host_grad);
GCTF_API int gctf_net_adam_step(gctf_net* net, float lr, float beta1, float beta2, float eps,
	float weight_decay);
GCTF_API int gctf_net_synchronize(gctf_net* net);

#ifdef __cplusplus
}
#endif
#endif /* GCTF_HIP_H */
