// gctf_host.cpp -- GCTFFind facade classes over the C-ABI.
#include "gctf_host.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

namespace GCTFFind {

CInput* CInput::GetInstance(void)
{
	static CInput s;
	return &s;
}

static gctf_ctx* s_ctx = nullptr;

gctf_ctx* SharedContext(void)
{
	if (s_ctx) return s_ctx;
	CInput* in = CInput::GetInstance();
	gctf_params p;
	p.kv = 300.0f; p.cs_mm = 2.7f; p.amp_contrast = 0.07f;
	p.ext_phase_deg = 0.0f; p.ext_phase_range_deg = 0.0f;
	p.tile_size = in->m_iTileSize;
	p.log_spect = in->m_iLogSpect;
	int rc = gctf_create(&p, in->m_iGpuID, &s_ctx);
	if (rc != GCTF_OK || !s_ctx) {
		char msg[128];
		snprintf(msg, sizeof msg, "gctf_create failed (status %d) on GPU %d", rc, in->m_iGpuID);
		throw std::runtime_error(msg);
	}
	return s_ctx;
}

const char* LastError(void) { return s_ctx ? gctf_error_string(s_ctx) : "no context"; }

static void require(int rc, const char* what)
{
	if (rc == GCTF_OK) return;
	std::string m = std::string(what) + ": " + LastError();
	throw std::runtime_error(m);
}

// ---- CCTFTheory --------------------------------------------------------
void CCTFTheory::Setup(float fKv, float fCs, float fAmpContrast, float fPixelSize, float fAstTol,
	float fExtPhase)
{
	m_fKv = fKv; m_fCs = fCs; m_fAmpContrast = fAmpContrast;
	m_fPixelSize = fPixelSize; m_fAstTol = fAstTol; m_fExtPhase = fExtPhase;
}

// ---- CRescaleImage -------------------------------------------------------
void CRescaleImage::Setup(int* piRawSize, float fRawPixSize)
{
	int sz[3];
	require(gctf_rescale_setup(SharedContext(), piRawSize[0], piRawSize[1], fRawPixSize, sz,
		&m_fPixSizeN), "CRescaleImage::Setup");
	m_aiNewSize[0] = sz[0]; m_aiNewSize[1] = sz[1];
	m_aiPadSizeN[0] = sz[2]; m_aiPadSizeN[1] = sz[1];
}

void CRescaleImage::DoIt(float* pfImage)
{
	require(gctf_rescale(SharedContext(), pfImage, 0), "CRescaleImage::DoIt");
}

void CRescaleImage::Prefetch(float* pfImage, int* piSize)
{
	require(gctf_rescale_prefetch(SharedContext(), pfImage, piSize[0], piSize[1]), "CRescaleImage::Prefetch");
}

float* CRescaleImage::GetScaledImg(void)
{
	float* p = nullptr;
	require(gctf_scaled_image(SharedContext(), &p), "CRescaleImage::GetScaledImg");
	return p;
}

// ---- CFindCtfBase ------------------------------------------------------
bool CFindCtfBase::mCheck(int rc, const char* what)
{
	if (rc == GCTF_OK) return true;
	fprintf(stderr, "%s: %s\n", what, LastError());
	return false;
}

void CFindCtfBase::mTake(const gctf_result& r)
{
	m_fDfMin = r.df_min;
	m_fDfMax = r.df_max;
	m_fAstAng = r.azimuth_deg;
	m_fExtPhase = r.ext_phase_deg;
	m_fScore = r.score;
	m_fCtfRes = r.ctf_res;
}

void CFindCtfBase::Setup1(CCTFTheory* t)
{
	require(gctf_setup1(SharedContext(), t->m_fKv, t->m_fCs, t->m_fAmpContrast, t->m_fPixelSize,
		t->m_fAstTol, t->m_fExtPhase), "CFindCtf2D::Setup1");
}

void CFindCtfBase::Setup2(int* piImgSize)
{
	require(gctf_setup2(SharedContext(), piImgSize[0], piImgSize[1]), "CFindCtfBase::Setup2");
}

void CFindCtfBase::SetPhase(float fInitPhase, float fPhaseRange)
{
	m_fExtPhase = fInitPhase;
	require(gctf_set_phase(SharedContext(), fInitPhase, fPhaseRange), "CFindCtfBase::SetPhase");
}

void CFindCtfBase::SetHalfSpect(float* pfCtfSpect)
{
	require(gctf_set_half_spect(SharedContext(), pfCtfSpect), "CFindCtfBase::SetHalfSpect");
}

float* CFindCtfBase::GetHalfSpect(bool bRaw, bool bToHost)
{
	// the device copy is internal to the library; both forms return a host copy
	(void)bToHost;
	int sz[2];
	gctf_spect_size(SharedContext(), sz, 1);
	float* out = new float[(size_t)sz[0] * sz[1]];
	if (!mCheck(gctf_get_half_spect(SharedContext(), bRaw ? 1 : 0, out), "GetHalfSpect")) {
		delete[] out;
		return nullptr;
	}
	return out;
}

void CFindCtfBase::GetSpectSize(int* piSize, bool bHalf)
{
	gctf_spect_size(SharedContext(), piSize, bHalf ? 1 : 0);
}

void CFindCtfBase::GenHalfSpectrum(float* gfPadImg)
{
	require(gctf_gen_half_spectrum(SharedContext(), gfPadImg), "CFindCtfBase::GenHalfSpectrum");
}

float* CFindCtfBase::GenFullSpectrum(void)
{
	int sz[2];
	gctf_spect_size(SharedContext(), sz, 0);
	float* out = new float[(size_t)sz[0] * sz[1]];
	if (!mCheck(gctf_full_spectrum(SharedContext(), out), "GenFullSpectrum")) {
		delete[] out;
		return nullptr;
	}
	return out;
}

bool CFindCtfBase::GenFullSpectrum(float* pfOut)
{
	return pfOut && mCheck(gctf_full_spectrum(SharedContext(), pfOut), "GenFullSpectrum");
}

void CFindCtfBase::ShowResult(void)
{
	printf("%9.2f  %9.2f  %6.2f  %6.2f  %8.5f\n\n", m_fDfMin, m_fDfMax, m_fAstAng, m_fExtPhase, m_fScore);
}

// ---- CFindCtf1D / CFindCtf2D --------------------------------------------
void CFindCtf1D::Do1D(void)
{
	gctf_result r;
	if (mCheck(gctf_do1d(SharedContext(), &r), "CFindCtf1D::Do1D")) mTake(r);
	else m_fScore = 0.0f;
	printf("1D estimate: %8.2f  %8.2f  %8.2f\n\n", m_fDfMin, m_fExtPhase, m_fScore);
}

void CFindCtf1D::Refine1D(float fInitDf, float fDfRange)
{
	gctf_result r;
	if (mCheck(gctf_refine1d(SharedContext(), fInitDf, fDfRange, &r), "CFindCtf1D::Refine1D")) mTake(r);
	else m_fScore = 0.0f;
	printf("1D estimate: %8.2f  %8.2f  %8.2f\n\n", m_fDfMin, m_fExtPhase, m_fScore);
}

void CFindCtf2D::Do2D(void)
{
	gctf_result r;
	if (mCheck(gctf_do2d(SharedContext(), &r), "CFindCtf2D::Do2D")) mTake(r);
	else m_fScore = 0.0f;
}

void CFindCtf2D::Refine(float afDfMean[2], float afAstRatio[2], float afAstAngle[2], float afExtPhase[2])
{
	gctf_result r;
	if (mCheck(gctf_refine(SharedContext(), afDfMean, afAstRatio, afAstAngle, afExtPhase, &r),
			"CFindCtf2D::Refine"))
		mTake(r);
	else
		m_fScore = 0.0f;
}

// ---- CFindCtfHelp (CFindCtfHelp.cpp:5-23) ---------------------------------
float CFindCtfHelp::CalcAstRatio(float fDfMin, float fDfMax)
{
	float fDiff = (fDfMax - fDfMin) * 0.5f;
	float fMean = (fDfMax + fDfMin) * 0.5f;
	return fMean <= 0.0f ? 0.0f : fDiff / fMean;
}

float CFindCtfHelp::CalcDfMin(float fDfMean, float fAstRatio) { return fDfMean * (1.0f - fAstRatio); }
float CFindCtfHelp::CalcDfMax(float fDfMean, float fAstRatio) { return fDfMean * (1.0f + fAstRatio); }

}  // namespace GCTFFind
