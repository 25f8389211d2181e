// gctf_host.hpp -- the reference's C++ call surface (FindCTF/CFindCTFInc.h)
// re-created on top of the C-ABI in include/gctf_hip.h.
//
// CRescaleImage, CCTFTheory, CFindCtfBase / CFindCtf1D / CFindCtf2D and
// CFindCtfHelp keep the reference's method names, argument meaning, units
// and public result fields, so CFindSeriesCtfs-style callers can switch to
// this library without changes to their logic.  Hidden inputs of the
// reference (CInput::m_iTileSize, m_iLogSpect, m_iGpuID) come from
// GCTFFind::CInput here as well.
#pragma once
#include <string>

#include "gctf_hip.h"

namespace GCTFFind {

// Process-wide run options, as the reference's CInput singleton
// (CMainInc.h:10-55).  Only the fields the CTF path reads live here.
class CInput {
public:
	static CInput* GetInstance(void);
	int m_iTileSize = 512;
	int m_iLogSpect = 0;
	int m_iGpuID = 0;
};

// One gctf_ctx per GPU shared by every facade object of the process.
gctf_ctx* SharedContext(void);
// Message of the last failed C-ABI call of the shared context.
const char* LastError(void);

class CCTFTheory {
public:
	// CCTFTheory::Setup (CFindCTFInc.h:36-43): kV, Cs (mm), amplitude
	// contrast, pixel size (A), astigmatism tolerance (A), extra phase (rad).
	void Setup(float fKv, float fCs, float fAmpContrast, float fPixelSize, float fAstTol,
		float fExtPhase);
	float GetPixelSize(void) const { return m_fPixelSize; }
	float m_fKv = 300.0f, m_fCs = 2.7f, m_fAmpContrast = 0.07f;
	float m_fPixelSize = 1.0f, m_fAstTol = 100.0f, m_fExtPhase = 0.0f;
};

class CRescaleImage {
public:
	// CRescaleImage::Setup(int* piRawSize, float fRawPixSize) (CFindCTFInc.h:428)
	void Setup(int* piRawSize, float fRawPixSize);
	// DoIt(float* pfImage): host image of the Setup size
	void DoIt(float* pfImage);
	// no reference counterpart: starts the device copy of the host image
	// that the next DoIt will rescale (piSize its nx, ny), so that the
	// transfer overlaps work in between; the image must stay unchanged
	// until that DoIt returns
	void Prefetch(float* pfImage, int* piSize);
	// device pointer, rows of m_aiPadSizeN[0] floats; do not free
	float* GetScaledImg(void);
	int m_aiNewSize[2] = {0, 0};
	int m_aiPadSizeN[2] = {0, 0};
	float m_fPixSizeN = 1.0f;
};

class CFindCtfBase {
public:
	virtual ~CFindCtfBase(void) {}
	void Setup1(CCTFTheory* pCtfTheory);
	void Setup2(int* piImgSize);
	void SetPhase(float fInitPhase, float fPhaseRange);   // degree
	void SetHalfSpect(float* pfCtfSpect);                 // host
	float* GetHalfSpect(bool bRaw, bool bToHost);         // new[] when bToHost
	void GetSpectSize(int* piSize, bool bHalf);
	void GenHalfSpectrum(float* gfPadImg);                // device image
	float* GenFullSpectrum(void);                         // new[], freed by caller
	// no reference counterpart: the same into the caller's buffer of
	// GetSpectSize(.., false) floats (e.g. page-locked memory, reused)
	bool GenFullSpectrum(float* pfOut);
	void ShowResult(void);
	float m_fDfMin = 0.0f;
	float m_fDfMax = 0.0f;
	float m_fAstAng = 0.0f;     // degree
	float m_fExtPhase = 0.0f;   // degree
	float m_fScore = 0.0f;
	float m_fCtfRes = 0.0f;     // A
protected:
	void mTake(const gctf_result& r);
	bool mCheck(int rc, const char* what);
};

class CFindCtf1D : public CFindCtfBase {
public:
	void Do1D(void);
	void Refine1D(float fInitDf, float fDfRange);
};

class CFindCtf2D : public CFindCtf1D {
public:
	void Do2D(void);
	// [0] = centre, [1] = full range (CFindCTFInc.h:606-611)
	void Refine(float afDfMean[2], float afAstRatio[2], float afAstAngle[2], float afExtPhase[2]);
};

class CFindCtfHelp {
public:
	static float CalcAstRatio(float fDfMin, float fDfMax);
	static float CalcDfMin(float fDfMean, float fAstRatio);
	static float CalcDfMax(float fDfMean, float fAstRatio);
};

}  // namespace GCTFFind
