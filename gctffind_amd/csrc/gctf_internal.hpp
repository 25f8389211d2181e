// gctf_internal.hpp -- shared device structures and kernel launchers of
// libgctf_hip.so.  Not part of the public C-ABI (include/gctf_hip.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gctf_fft.hpp"

namespace gctf {

// Candidate-stage kinds of the 2D search (CFindDefocus2D.cpp:153-361).
enum StageKind : int {
	ST_NONE = 0,
	ST_INIT = 2,     // mIterate's first mCorrelate
	ST_ASTIG = 3,    // mFindAstig grid
	ST_ASTMAG = 4,   // mRefineAstMag
	ST_ASTANG = 5,   // mRefineAstAng
	ST_DFMEAN = 6,   // mRefineDfMean
	ST_PHASE = 7,    // mRefinePhase
};

// Largest candidate count of one stage: the 51 x 51 astigmatism grid.
constexpr int kMaxCands = 2601;
// Pixel chunks the 2D correlation is split into (partial sums per chunk).
constexpr int kPixChunks = 8;

// Per-plan CTF constants in pixel units (CCTFTheory.cpp:77-78,
// GCalcCTF2D.cu:88-97).
struct CtfConsts {
	float wl;         // wavelength / pixel
	float cs;         // Cs / pixel
	float amp_phase;  // atanf(ac / sqrtf(1 - ac^2))
	float pix;        // working pixel size (A)
};

// Candidate counts of the 2D search program.  The plain program is the
// reference's (CFindDefocus2D.cpp:153-361): a 51 x 51 astigmatism grid, then
// 19 rounds of 21 / 51 / 31 candidates, defocus step floor 20 A.  The
// learned path runs a smaller program inside the proposal's window and
// ends with a fine defocus stage.
struct SearchProgram {
	int n_ast, n_ang;     // mFindAstig grid
	int n_mag, n_angl;    // mRefineAstMag / mRefineAstAng candidates
	int n_df;             // mRefineDfMean candidates
	int iters;            // mIterate's iIterations
	float df_minstep;     // mRefineDfMean step floor (A)
	int fine_n;           // final defocus stage (0: none)
	float fine_step;      // its step (A)
};

__host__ __device__ constexpr SearchProgram plain_program()
{
	return SearchProgram{51, 51, 21, 51, 31, 20, 20.0f, 0, 0.0f};
}

// Per-micrograph search state; mirrors the fields of CFindCtfBase,
// CFindDefocus1D and CFindDefocus2D (CFindCTFInc.h:480-562).
struct SearchState {
	// CFindCtfBase outputs (degrees, A)
	float df_min, df_max, ast_ang, ext_phase, score, ctf_res;
	// CFindCtfBase::SetPhase range (degrees); ext_phase above holds the
	// current CFindCtfBase::m_fExtPhase
	float phase_range;
	// CFindDefocus2D
	float dfmean, ast, ang, phase, ccmax, oldcc;
	float dfrange[2], astrange[2], angrange[2], phrange[2];
	float tot_df, tot_ast, tot_ang;
	int done;
	int iter;
	// pending candidate stage
	int st_kind;
	int st_n;        // number of candidates
	int st_nast;     // astig grid: candidates per angle row
	float st_a0, st_astep;   // first value / step (ast for ASTIG)
	float st_b0, st_bstep;   // angle grid for ASTIG
	float st_lo, st_hi;      // phase window for ST_PHASE
	float st_old;            // mRefineDfMean: the df before the stage
	int lowpassed;           // a12 gate outcome
	SearchProgram pg;        // candidate counts of the running 2D program
	int skip;                // 1: the image sits out the running search pass
	int fallback;            // learned path: 1 when the plain search must decide
	float conf;              // learned path: proposal confidence
	// pending 1D stage (CFindDefocus1D::mBrutalForceSearch): candidate
	// i = iph * s1_ndf + idf at defocus s1_df0 + idf s1_dfstep A and phase
	// s1_ph0 + iph s1_phstep degrees (clamped to [0, 150]); s1_stage 0 is
	// CFindCtf1D::mFindDefocus, 1 mRefineDefocus
	float s1_df0, s1_dfstep, s1_ph0, s1_phstep;
	int s1_ndf, s1_n, s1_stage;
};

// Per-plan 2D-ring tables in structure-of-arrays form (device pointers).
struct RingTables {
	int n;                 // ring pixel count
	const int* idx;        // y * cmpx + x
	const float* a;        // (pi'/pi) * wl * s^2          (chi/pi per unit df)
	const float* b;        // (pi'/pi) * 0.5 * wl^3 * s^4 * Cs
	const float* c2;       // cos(2 theta)
	const float* s2;       // sin(2 theta)
	const float* w;        // exp(-16 s^2)
};

// Per-micrograph search plan (device array indexed by image): the CTF
// constants at the micrograph's working pixel size, its correlation ring and
// its 1D bin range.  Micrographs of one batch may differ in all three.
struct ImgPlan {
	CtfConsts cc;
	RingTables ring;
	float lo1d, hi1d;
};

}  // namespace gctf
