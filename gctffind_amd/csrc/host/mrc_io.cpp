// mrc_io.cpp -- see mrc_io.hpp.
#include "mrc_io.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>

namespace GCTFFind {

static uint32_t bswap32(uint32_t v)
{
	return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
}

static uint16_t bswap16(uint16_t v) { return (uint16_t)((v >> 8) | (v << 8)); }

int MrcModeBytes(int mode)
{
	switch (mode) {
	case 0: case 5: return 1;
	case 1: case 6: case 12: return 2;
	case 2: return 4;
	default: return 0;
	}
}

static float half_to_float(uint16_t h)
{
	uint32_t s = (h >> 15) & 1u, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
	float v;
	if (e == 0) v = std::ldexp((float)m, -24);
	else if (e == 31) v = m ? NAN : INFINITY;
	else v = std::ldexp((float)(m | 0x400u), (int)e - 25);
	return s ? -v : v;
}

bool MrcReader::Open(const std::string& path)
{
	Close();
	m_fp = fopen(path.c_str(), "rb");
	if (!m_fp) return false;
	unsigned char hdr[1024];
	if (fread(hdr, 1, 1024, m_fp) != 1024) { Close(); return false; }
	int32_t w[4];
	memcpy(w, hdr, 16);
	// little-endian unless nx/ny/nz/mode only make sense swapped
	auto plausible = [](const int32_t* v) {
		return v[0] > 0 && v[1] > 0 && v[2] > 0 && v[0] < (1 << 20) && v[1] < (1 << 20) && v[3] >= 0 && v[3] < 32;
	};
	bool swap = false;
	if (!plausible(w)) {
		int32_t s[4];
		for (int i = 0; i < 4; ++i) s[i] = (int32_t)bswap32((uint32_t)w[i]);
		if (!plausible(s)) { Close(); return false; }
		memcpy(w, s, 16);
		swap = true;
	}
	m_h.nx = w[0]; m_h.ny = w[1]; m_h.nz = w[2]; m_h.mode = w[3];
	int32_t ns;
	memcpy(&ns, hdr + 92, 4);
	m_h.nsymbt = swap ? (int32_t)bswap32((uint32_t)ns) : ns;
	int16_t ni, nf;
	memcpy(&ni, hdr + 128, 2);
	memcpy(&nf, hdr + 130, 2);
	if (swap) { ni = (int16_t)bswap16((uint16_t)ni); nf = (int16_t)bswap16((uint16_t)nf); }
	m_h.numintegers = ni; m_h.numfloats = nf;
	m_h.swapped = swap;
	if (MrcModeBytes(m_h.mode) == 0 || m_h.nsymbt < 0) { Close(); return false; }
	// the header's sections must fit the file: a header that promises more
	// data than the file holds is refused here, before any section buffer
	// is sized from it
	if (fseeko(m_fp, 0, SEEK_END) != 0) { Close(); return false; }
	const long long fsize = (long long)ftello(m_fp);
	const long long need = 1024LL + (long long)m_h.nsymbt
		+ (long long)m_h.nz * (long long)m_h.nx * (long long)m_h.ny * MrcModeBytes(m_h.mode);
	if (fsize < 0 || need > fsize) { Close(); return false; }
	return true;
}

void MrcReader::Close(void)
{
	if (m_fp) fclose(m_fp);
	m_fp = nullptr;
}

bool MrcReader::ReadSection(int iz, std::vector<float>& out)
{
	if (!m_fp || iz < 0 || iz >= m_h.nz) return false;
	out.resize((size_t)m_h.nx * m_h.ny);
	return ReadSection(iz, out.data());
}

// The section's bytes are read into the tail of `out` and converted to
// float32 front to back in place: value i (bpp bytes at byte offset
// n (4 - bpp) + bpp i) is read before float i (bytes [4 i, 4 i + 4)) is
// written, and float i overwrites only values < i + 1, since
// 4 (i + 1) <= n (4 - bpp) + bpp (i + 1) for i < n.
bool MrcReader::ReadSection(int iz, float* out)
{
	if (!m_fp || !out || iz < 0 || iz >= m_h.nz) return false;
	const int bpp = MrcModeBytes(m_h.mode);
	if (bpp == 0) return false;
	const size_t n = (size_t)m_h.nx * m_h.ny;
	const long long off = 1024LL + m_h.nsymbt + (long long)iz * (long long)n * bpp;
	if (fseeko(m_fp, off, SEEK_SET) != 0) return false;
	unsigned char* raw = reinterpret_cast<unsigned char*>(out) + n * (4 - bpp);
	if (fread(raw, 1, n * bpp, m_fp) != n * bpp) return false;
	const bool sw = m_h.swapped;
	switch (m_h.mode) {
	case 0: case 5:
		for (size_t i = 0; i < n; ++i) out[i] = (float)raw[i];
		break;
	case 1:
		for (size_t i = 0; i < n; ++i) {
			uint16_t v; memcpy(&v, raw + 2 * i, 2);
			if (sw) v = bswap16(v);
			out[i] = (float)(int16_t)v;
		}
		break;
	case 6:
		for (size_t i = 0; i < n; ++i) {
			uint16_t v; memcpy(&v, raw + 2 * i, 2);
			if (sw) v = bswap16(v);
			out[i] = (float)v;
		}
		break;
	case 12:
		for (size_t i = 0; i < n; ++i) {
			uint16_t v; memcpy(&v, raw + 2 * i, 2);
			if (sw) v = bswap16(v);
			out[i] = half_to_float(v);
		}
		break;
	case 2:
		if (sw)
			for (size_t i = 0; i < n; ++i) {
				uint32_t v; memcpy(&v, raw + 4 * i, 4);
				v = bswap32(v);
				memcpy(&out[i], &v, 4);
			}
		break;
	default:
		return false;
	}
	return true;
}

bool MrcReader::Tilt(int iz, float* tilt)
{
	if (!m_fp || m_h.nsymbt <= 0 || m_h.numfloats < 1) return false;
	const int rec = 4 * (m_h.numintegers + m_h.numfloats);
	if (rec <= 0 || (long long)rec * (iz + 1) > m_h.nsymbt) return false;
	if (fseeko(m_fp, 1024LL + (long long)rec * iz + 4LL * m_h.numintegers, SEEK_SET) != 0) return false;
	uint32_t v;
	if (fread(&v, 4, 1, m_fp) != 1) return false;
	if (m_h.swapped) v = bswap32(v);
	memcpy(tilt, &v, 4);
	return true;
}

bool MrcWriter::Open(const std::string& path, int nx, int ny, int nz, float pix)
{
	Close();
	m_fp = fopen(path.c_str(), "wb+");
	if (!m_fp) return false;
	m_nx = nx; m_ny = ny; m_nz = nz; m_pix = pix;
	m_sum = 0.0; m_count = 0; m_min = INFINITY; m_max = -INFINITY;
	std::vector<unsigned char> zero(1024, 0);
	return fwrite(zero.data(), 1, 1024, m_fp) == 1024;
}

bool MrcWriter::WriteSection(int iz, const float* data)
{
	if (!m_fp || iz < 0 || iz >= m_nz) return false;
	const size_t n = (size_t)m_nx * m_ny;
	if (fseeko(m_fp, 1024LL + (long long)iz * (long long)n * 4, SEEK_SET) != 0) return false;
	if (fwrite(data, 4, n, m_fp) != n) return false;
	for (size_t i = 0; i < n; ++i) {
		m_sum += data[i];
		if (data[i] < m_min) m_min = data[i];
		if (data[i] > m_max) m_max = data[i];
	}
	m_count += (long long)n;
	return true;
}

bool MrcWriter::Close(void)
{
	if (!m_fp) return true;
	unsigned char h[1024];
	memset(h, 0, sizeof h);
	int32_t iv[10] = {m_nx, m_ny, m_nz, 2, 0, 0, 0, m_nx, m_ny, m_nz};
	memcpy(h, iv, sizeof iv);
	float cell[6] = {m_nx * m_pix, m_ny * m_pix, m_nz * m_pix, 90.f, 90.f, 90.f};
	memcpy(h + 40, cell, sizeof cell);
	int32_t axes[3] = {1, 2, 3};
	memcpy(h + 64, axes, sizeof axes);
	float st[3] = {m_count ? m_min : 0.f, m_count ? m_max : 0.f,
		m_count ? (float)(m_sum / (double)m_count) : 0.f};
	memcpy(h + 76, st, sizeof st);
	memcpy(h + 208, "MAP ", 4);
	h[212] = 0x44; h[213] = 0x44;
	bool ok = fseeko(m_fp, 0, SEEK_SET) == 0 && fwrite(h, 1, 1024, m_fp) == 1024;
	ok = (fclose(m_fp) == 0) && ok;
	m_fp = nullptr;
	return ok;
}

}  // namespace GCTFFind
