// mrc_io.hpp -- MRC2014 stack reading and float32 writing (public format:
// 1024-byte main header, `nsymbt` bytes of extended header, then sections).
// Modes read: 0 and 5 (8-bit, read unsigned as the reference does,
// CLoadImages.cpp:149-155), 1 (int16), 2 (float32), 6 (uint16) and 12
// (float16).  Per-section tilt angles come from an extended header of
// `numintegers` ints followed by `numfloats` floats per section (header bytes
// 129-132); the tilt is the first float.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace GCTFFind {

struct MrcHeader {
	int nx = 0, ny = 0, nz = 0, mode = 2;
	int nsymbt = 0;
	int numintegers = 0, numfloats = 0;
	bool swapped = false;
};

class MrcReader {
public:
	bool Open(const std::string& path);
	void Close(void);
	const MrcHeader& Header(void) const { return m_h; }
	// section iz converted to float32 (nx*ny values); false on I/O error or
	// an unsupported mode
	bool ReadSection(int iz, std::vector<float>& out);
	// the same into out, which holds at least nx*ny floats
	bool ReadSection(int iz, float* out);
	// tilt of section iz from the extended header; false when there is none
	bool Tilt(int iz, float* tilt);
	~MrcReader() { Close(); }
private:
	FILE* m_fp = nullptr;
	MrcHeader m_h;
};

// Writes a float32 (mode 2) stack of nz sections of nx*ny, pixel size pix.
class MrcWriter {
public:
	bool Open(const std::string& path, int nx, int ny, int nz, float pix);
	bool WriteSection(int iz, const float* data);
	bool Close(void);
	~MrcWriter() { Close(); }
private:
	FILE* m_fp = nullptr;
	int m_nx = 0, m_ny = 0, m_nz = 0;
	float m_pix = 1.0f;
	double m_sum = 0.0;
	float m_min = 0.f, m_max = 0.f;
	long long m_count = 0;
};

int MrcModeBytes(int mode);

}  // namespace GCTFFind
