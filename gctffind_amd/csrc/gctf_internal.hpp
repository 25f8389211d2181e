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
	int pad_;
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

}  // namespace gctf
