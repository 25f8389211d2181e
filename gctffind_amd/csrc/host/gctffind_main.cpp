// gctffind_main.cpp -- command-line front end with the reference's flags,
// defaults and outputs (CInput.cpp:25-179, CProcessMain.cpp:16-39,
// CFindSeriesCtfs.cpp:50-239, CSaveCtfResults.cpp:34-117), driving the
// GCTFFind facade classes of gctf_host.hpp on one GPU.
#include <dirent.h>
#include <strings.h>
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "gctf_host.hpp"
#include "mrc_io.hpp"

using namespace GCTFFind;

namespace {

struct Options {
	std::string in_mrc, out_mrc, out_ctf, ang_file, in_suffix;
	float kv = 300.0f, cs = 2.7f, ac = 0.07f, pix = 1.0f;
	float ext_phase[2] = {0.0f, 0.0f};
	float tilt_range[2] = {0.0f, 0.0f};
	int tile = 512, log_spect = 0, serial = 0, gpu = 0;
};

// CParseArgs::FindVals (Util/CParseArgs.cpp:25-52): case-insensitive tag;
// the values are the run of numeric tokens after it.
struct Args {
	int argc;
	char** argv;
	int find(const char* tag, int* count) const
	{
		*count = 0;
		int at = -1;
		for (int i = 1; i < argc; ++i)
			if (strcasecmp(argv[i], tag) == 0) { at = i + 1; break; }
		if (at < 0) return -1;
		for (int j = at; j < argc; ++j) {
			char* end = nullptr;
			strtod(argv[j], &end);
			if (end != argv[j] && *end == '\0') { ++*count; continue; }
			strtol(argv[j], &end, 0);
			if (end != argv[j] && *end == '\0') { ++*count; continue; }
			break;
		}
		return at;
	}
	void str(const char* tag, std::string& out) const
	{
		int n, at = find(tag, &n);
		if (at > 0 && at < argc) out = argv[at];
	}
	void floats(const char* tag, float* v, int maxn) const
	{
		int n, at = find(tag, &n);
		if (at < 0) return;
		n = std::min(n, maxn);
		for (int i = 0; i < n && at + i < argc; ++i) sscanf(argv[at + i], "%f", v + i);
	}
	void ints(const char* tag, int* v, int maxn) const
	{
		int n, at = find(tag, &n);
		if (at < 0) return;
		n = std::min(n, maxn);
		for (int i = 0; i < n && at + i < argc; ++i) sscanf(argv[at + i], "%d", v + i);
	}
};

void show_tags()
{
	printf("-InMrc         input MRC file (single image or a stack of tilt images)\n"
	       "-OutMrc        output MRC file of the averaged spectra with the fitted CTF\n"
	       "-OutCtf        output text file of the CTF parameters, one line per image\n"
	       "-AngFile       text file with one tilt angle per line\n"
	       "-TiltRange     min and max tilt of a fixed-step tilt series\n"
	       "-kV            high tension in kV (300)\n"
	       "-Cs            spherical aberration in mm (2.7)\n"
	       "-AmpContrast   amplitude contrast (0.07)\n"
	       "-PixSize       pixel size in A (1.0)\n"
	       "-ExtPhase      extra phase and search range in degree (0 0)\n"
	       "-TileSize      tile size (512)\n"
	       "-LogSpect      1: box mean instead of mean of square roots (0)\n"
	       "-Serial        1: every file in -InMrc's folder with -InSuffix\n"
	       "-InSuffix      suffix of the files to process in serial mode\n"
	       "-Gpu           GPU id (0)\n");
}

void parse(int argc, char** argv, Options& o)
{
	Args a{argc, argv};
	a.str("-InMrc", o.in_mrc);
	a.str("-OutMrc", o.out_mrc);
	a.str("-OutCtf", o.out_ctf);
	a.str("-InSuffix", o.in_suffix);
	a.str("-AngFile", o.ang_file);
	a.floats("-kV", &o.kv, 1);
	a.floats("-Cs", &o.cs, 1);
	a.floats("-AmpContrast", &o.ac, 1);
	a.floats("-PixSize", &o.pix, 1);
	a.floats("-ExtPhase", o.ext_phase, 2);
	if (o.ext_phase[1] < 0) o.ext_phase[1] = 0.0f;
	a.floats("-TiltRange", o.tilt_range, 2);
	o.tilt_range[0] = fmaxf(o.tilt_range[0], -70.1f);
	o.tilt_range[1] = fminf(o.tilt_range[1], 70.1f);
	a.ints("-TileSize", &o.tile, 1);
	a.ints("-LogSpect", &o.log_spect, 1);
	a.ints("-Serial", &o.serial, 1);
	a.ints("-Gpu", &o.gpu, 1);
	printf("\n%-15s  %s\n%-15s  %s\n%-15s  %s\n%-15s  %s\n%-15s  %s\n", "-InMrc", o.in_mrc.c_str(),
		"-InSuffix", o.in_suffix.c_str(), "-OutMrc", o.out_mrc.c_str(), "-OutCtf", o.out_ctf.c_str(),
		"-AngFile", o.ang_file.c_str());
	printf("%-15s  %f\n%-15s  %f\n%-15s  %f\n%-15s  %f\n", "-kV", o.kv, "-Cs", o.cs, "-AmpContrast", o.ac,
		"-PixSize", o.pix);
	printf("%-15s  %f  %f\n%-15s  %d\n%-15s  %.2f  %.2f\n%-15s  %d\n%-15s  %d\n%-15s  %d\n\n", "-ExtPhase",
		o.ext_phase[0], o.ext_phase[1], "-TileSize", o.tile, "-TiltRange", o.tilt_range[0],
		o.tilt_range[1], "-LogSpect", o.log_spect, "-Serial", o.serial, "-Gpu", o.gpu);
}

struct Package {          // CCtfPackage (CMainInc.h:57-80)
	std::string file;
	int idx = -1;
	int section = 0;
	float tilt = 0.0f;
	float df_min = 0, df_max = 0, azimuth = 0, ext_phase = 0, score = 0;
};

// CLoadAngFile::DoIt (CLoadAngFile.cpp:44-72)
std::vector<float> load_angles(const std::string& f)
{
	std::vector<float> t;
	if (f.empty()) return t;
	FILE* fp = fopen(f.c_str(), "r");
	if (!fp) return t;
	char line[256];
	while (fgets(line, sizeof line, fp)) {
		if (line[0] == '#') continue;
		float v;
		if (sscanf(line, "%f", &v) == 1 && fabsf(v) < 85) t.push_back(v);
		if (t.size() >= 1024) break;
	}
	fclose(fp);
	return t;
}

// CInputFolder::FindZeroTilt (CInputFolder.cpp:49-68)
int find_zero_tilt(const std::vector<Package>& p)
{
	if (p.size() < 2) return -1;
	if (fabsf(p[0].tilt - p[1].tilt) < 0.0001f) return (int)p.size() / 2;
	int z = -1;
	float m = 10000.0f;
	for (size_t i = 0; i < p.size(); ++i) {
		float t = fabsf(p[i].tilt);
		if (t >= m) continue;
		m = t; z = (int)i;
	}
	return z;
}

std::string base_name(const std::string& p)
{
	size_t s = p.rfind('/');
	return s == std::string::npos ? p : p.substr(s + 1);
}

// CSaveCtfResults::SaveCTF / SaveImod (CSaveCtfResults.cpp:34-117)
void save_ctf(const Options& o, const std::vector<Package>& p)
{
	if (o.out_ctf.empty()) return;
	FILE* fp = fopen(o.out_ctf.c_str(), "w");
	if (!fp) {
		printf("CTF file: %s\n   cannot be opened, CTF results not saved.\n\n", o.out_ctf.c_str());
		return;
	}
	fprintf(fp, "# Idx   tilt     dfMin      dfMax   azimuth extPhase  score    fileName\n");
	for (const auto& k : p)
		fprintf(fp, "%4d  %6.2f %10.2f %10.2f %7.2f %7.2f %8.4f  %s\n", k.idx, k.tilt, k.df_min, k.df_max,
			k.azimuth, k.ext_phase, k.score, k.file.c_str());
	fclose(fp);
}

void save_imod(const Options& o, const std::vector<Package>& p, bool tomo)
{
	if (o.out_ctf.empty() || !tomo || p.empty()) return;
	std::string f = o.out_ctf;
	size_t dot = f.rfind('.');
	if (dot == std::string::npos) f += "_Imod.txt";
	else f = f.substr(0, dot) + "_Imod.txt";
	FILE* fp = fopen(f.c_str(), "w");
	if (!fp) return;
	const bool ph = p[0].ext_phase != 0;
	fprintf(fp, ph ? "5  0  0.0  0.0  0.0  3\n" : "1  0  0.0  0.0  0.0  3\n");
	for (size_t i = 0; i < p.size(); ++i) {
		const Package& k = p[i];
		if (ph)
			fprintf(fp, "%4d  %4d  %7.2f  %7.2f  %8.2f  %8.2f  %7.2f  %8.2f\n", (int)i + 1, (int)i + 1, k.tilt,
				k.tilt, k.df_min * 0.1f, k.df_max * 0.1f, k.azimuth, k.ext_phase);
		else
			fprintf(fp, "%4d  %4d  %7.2f  %7.2f  %8.2f  %8.2f  %7.2f\n", (int)i + 1, (int)i + 1, k.tilt,
				k.tilt, k.df_min * 0.1f, k.df_max * 0.1f, k.azimuth);
	}
	fclose(fp);
}

// CLoadImages::ThreadMain (CLoadImages.cpp:64-137): micrographs are read
// ahead of the GPU, in processing order, while it works on earlier ones.
// `readers` threads read different files at once, each straight into a
// buffer of a pool of page-locked host memory (gctf_host_alloc; ordinary
// memory if that fails).  The consumer hands a buffer back once the GPU has
// copied its image, so at most `ahead` micrographs wait and no file costs a
// fresh allocation.
class ImageLoader {
public:
	struct Item {
		bool ok = false;
		int nx = 0, ny = 0, mode = 0;
		float* px = nullptr;   // nx*ny floats, the loader's until give_back
		int buf = -1;
	};
	ImageLoader(const std::vector<Package>& pk, const std::vector<int>& order, size_t ahead, int readers)
		: pk_(pk), order_(order), ahead_(std::max<size_t>(ahead, 1)), items_(order.size()),
		  ready_(order.size(), 0), bufs_(ahead_ + 1)
	{
		for (int i = 0; i < (int)bufs_.size(); ++i) free_.push_back(i);
		for (int t = 0; t < std::max(readers, 1); ++t) th_.emplace_back([this] { run(); });
	}
	~ImageLoader()
	{
		{
			std::lock_guard<std::mutex> l(m_);
			stop_ = true;
		}
		cv_.notify_all();
		for (auto& t : th_) t.join();
		for (auto& b : bufs_) release(b);
	}
	// The k-th micrograph of the processing order; waits until it is read.
	Item take(size_t k)
	{
		std::unique_lock<std::mutex> l(m_);
		cv_.wait(l, [&] { return ready_[k] != 0; });
		return items_[k];
	}
	// The k-th micrograph if it has been read, without waiting.
	bool peek(size_t k, Item* out)
	{
		std::lock_guard<std::mutex> l(m_);
		if (k >= ready_.size() || !ready_[k]) return false;
		*out = items_[k];
		return true;
	}
	// Returns the buffer of the k-th micrograph (taken with take(k)); no
	// copy from it may be pending.
	void give_back(size_t k, const Item& it)
	{
		{
			std::lock_guard<std::mutex> l(m_);
			if (it.buf >= 0) free_.push_back(it.buf);
			taken_ = k + 1;
		}
		cv_.notify_all();
	}

private:
	struct Buf {
		float* p = nullptr;
		size_t cap = 0;
		bool pinned = false;
	};
	static void release(Buf& b)
	{
		if (b.p) {
			if (b.pinned) gctf_host_free(b.p);
			else std::free(b.p);
		}
		b = Buf();
	}
	static bool reserve(Buf& b, size_t n)
	{
		if (b.cap >= n) return true;
		release(b);
		void* p = nullptr;
		if (gctf_host_alloc(n * sizeof(float), &p) == GCTF_OK) {
			b.pinned = true;
		} else {
			p = std::malloc(n * sizeof(float));
			if (!p) return false;
		}
		b.p = static_cast<float*>(p);
		b.cap = n;
		return true;
	}
	void run()
	{
		for (;;) {
			size_t k;
			int bi;
			{
				std::unique_lock<std::mutex> l(m_);
				cv_.wait(l, [&] {
					return stop_ || next_ >= order_.size() || (next_ < taken_ + ahead_ && !free_.empty());
				});
				if (stop_ || next_ >= order_.size()) return;
				k = next_++;
				bi = free_.back();
				free_.pop_back();
			}
			const Package& p = pk_[order_[k]];
			Item it;
			it.buf = bi;
			MrcReader r;
			// a failed read, including a buffer that cannot be allocated, is
			// reported to the consumer as a read error
			if (r.Open(p.file)) {
				it.nx = r.Header().nx;
				it.ny = r.Header().ny;
				it.mode = r.Header().mode;
				Buf& b = bufs_[bi];
				it.ok = reserve(b, (size_t)it.nx * it.ny) && r.ReadSection(p.section, b.p);
				it.px = b.p;
			}
			{
				std::lock_guard<std::mutex> l(m_);
				items_[k] = it;
				ready_[k] = 1;
			}
			cv_.notify_all();
		}
	}
	const std::vector<Package>& pk_;
	const std::vector<int>& order_;
	const size_t ahead_;
	std::vector<Item> items_;
	std::vector<char> ready_;
	std::vector<Buf> bufs_;
	std::vector<int> free_;
	size_t next_ = 0, taken_ = 0;
	bool stop_ = false;
	std::mutex m_;
	std::condition_variable cv_;
	std::vector<std::thread> th_;
};

// CAsyncSaveImages (CAsyncSaveImages.cpp:31-110): full spectra are written on
// a thread of their own, in the order they are handed over.  Stack sections
// go to one file opened at the first; serial inputs get one file each.
class SpectrumWriter {
public:
	struct Job {
		std::string file;
		bool stack = false;
		int section = 0, nz = 1, nx = 0, ny = 0;
		std::vector<float> px;
	};
	SpectrumWriter() { th_ = std::thread([this] { run(); }); }
	~SpectrumWriter() { finish(); }
	void push(Job j)
	{
		{
			std::lock_guard<std::mutex> l(m_);
			q_.push_back(std::move(j));
		}
		cv_.notify_one();
	}
	// A buffer of n floats for a job: one the writer has finished with when
	// there is one, so that no spectrum costs a fresh allocation and its
	// page faults.
	std::vector<float> buffer(size_t n)
	{
		std::vector<float> v;
		{
			std::lock_guard<std::mutex> l(m_);
			if (!free_.empty()) {
				v = std::move(free_.back());
				free_.pop_back();
			}
		}
		v.resize(n);
		return v;
	}
	// Writes what is queued, then closes the stack file.
	void finish()
	{
		if (!th_.joinable()) return;
		{
			std::lock_guard<std::mutex> l(m_);
			done_ = true;
		}
		cv_.notify_one();
		th_.join();
	}

private:
	void run()
	{
		std::unique_ptr<MrcWriter> stack;
		for (;;) {
			Job j;
			{
				std::unique_lock<std::mutex> l(m_);
				cv_.wait(l, [&] { return done_ || !q_.empty(); });
				if (q_.empty()) break;
				j = std::move(q_.front());
				q_.pop_front();
			}
			if (j.stack) {
				if (!stack) {
					stack = std::make_unique<MrcWriter>();
					if (!stack->Open(j.file, j.nx, j.ny, j.nz, 1.0f)) fprintf(stderr, "cannot write %s\n", j.file.c_str());
				}
				stack->WriteSection(j.section, j.px.data());
			} else {
				MrcWriter w;
				if (w.Open(j.file, j.nx, j.ny, 1, 1.0f)) { w.WriteSection(0, j.px.data()); w.Close(); }
			}
			std::lock_guard<std::mutex> l(m_);
			if (free_.size() < 8) free_.push_back(std::move(j.px));
		}
		if (stack) stack->Close();
	}
	std::deque<Job> q_;
	std::vector<std::vector<float>> free_;   // buffers of written jobs, for reuse
	bool done_ = false;
	std::mutex m_;
	std::condition_variable cv_;
	std::thread th_;
};

// CAsyncSaveImages::mEmbedSerial (CAsyncSaveImages.cpp:112-132)
std::string serial_out(const std::string& out, const std::string& serial)
{
	std::string f = out;
	size_t d = f.find(".mrc");
	if (d == std::string::npos) return f + serial + ".mrc";
	f = f.substr(0, d) + serial;
	if (serial.find("mrc") == std::string::npos) f += ".mrc";
	return f;
}

}  // namespace

int main(int argc, char** argv)
{
	if (argc == 1) { printf("Use GCtfFind --help to get more information.\n"); return 0; }
	if (strstr(argv[1], "--help")) { show_tags(); return 0; }
	if (strstr(argv[1], "--version")) { printf("GCtfFind (MI355X) 1.1.2-hip\n"); return 0; }
	if (strcmp(argv[1], "--mrcinfo") == 0 && argc > 2) {
		// header, per-section mean and tilt of an MRC file (no GPU use)
		MrcReader r;
		if (!r.Open(argv[2])) { printf("error\n"); return 1; }
		const MrcHeader& h = r.Header();
		printf("nx %d ny %d nz %d mode %d nsymbt %d nint %d nfloat %d swapped %d\n", h.nx, h.ny, h.nz, h.mode,
			h.nsymbt, h.numintegers, h.numfloats, h.swapped ? 1 : 0);
		std::vector<float> v;
		for (int z = 0; z < h.nz; ++z) {
			if (!r.ReadSection(z, v)) { printf("error section %d\n", z); return 1; }
			double s = 0.0;
			for (float f : v) s += f;
			float t = 0.0f;
			int ht = r.Tilt(z, &t) ? 1 : 0;
			printf("section %d mean %.9g first %.9g last %.9g tilt %d %.6g\n", z, s / v.size(), v.front(),
				v.back(), ht, t);
		}
		if (argc > 3) {   // rewrite as a float32 stack
			MrcWriter w;
			if (!w.Open(argv[3], h.nx, h.ny, h.nz, 1.0f)) return 1;
			for (int z = 0; z < h.nz; ++z) { r.ReadSection(z, v); w.WriteSection(z, v.data()); }
			if (!w.Close()) return 1;
		}
		return 0;
	}
	auto t_start = std::chrono::steady_clock::now();
	Options o;
	parse(argc, argv, o);
	if (strcasecmp(o.in_mrc.c_str(), o.out_mrc.c_str()) == 0 || strcasecmp(o.in_mrc.c_str(), o.out_ctf.c_str()) == 0) {
		fprintf(stderr, "Error: input and output files have the same name,\n\n");
		return 1;
	}
	CInput* in = CInput::GetInstance();
	in->m_iTileSize = o.tile;
	in->m_iLogSpect = o.log_spect;
	in->m_iGpuID = o.gpu;
	try {
		SharedContext();
	} catch (const std::exception& e) {
		printf("Error: Invalid GPU (%d): %s\n\n", o.gpu, e.what());
		return 1;
	}

	// ---- packages (CInputFolder::ReadFiles, CInputFolder.cpp:131-270)
	std::vector<Package> pk;
	bool tomo = false;
	std::string prefix_dir, prefix, suffix = o.in_suffix;
	if (o.serial == 0) {
		MrcReader r;
		if (!r.Open(o.in_mrc)) { fprintf(stderr, "Error: cannot open %s\n", o.in_mrc.c_str()); return 1; }
		const int nz = r.Header().nz;
		for (int i = 0; i < nz; ++i) {
			Package p;
			p.file = o.in_mrc; p.section = i;
			if (r.Header().numfloats > 1) r.Tilt(i, &p.tilt);
			pk.push_back(p);
		}
		tomo = nz > 1;
		if (tomo) {
			std::vector<float> ang = load_angles(o.ang_file);
			if ((int)ang.size() == nz) {
				for (int i = 0; i < nz; ++i) pk[i].tilt = ang[i];
			} else if (fabsf(o.tilt_range[1] - o.tilt_range[0]) >= 10.0f) {
				float step = (o.tilt_range[1] - o.tilt_range[0]) / (nz - 1);
				for (int i = 0; i < nz; ++i) pk[i].tilt = o.tilt_range[0] + i * step;
			}
		}
	} else {
		size_t s = o.in_mrc.rfind('/');
		prefix_dir = s == std::string::npos ? "./" : o.in_mrc.substr(0, s + 1);
		prefix = s == std::string::npos ? o.in_mrc : o.in_mrc.substr(s + 1);
		DIR* d = opendir(prefix_dir.c_str());
		if (!d) { fprintf(stderr, "Error: cannot open folder\n   %s\n", prefix_dir.c_str()); return 1; }
		std::vector<std::string> names;
		while (dirent* e = readdir(d)) {
			std::string n = e->d_name;
			if (n.empty() || n[0] == '.') continue;
			if (!prefix.empty() && n.find(prefix) == std::string::npos) continue;
			if (!suffix.empty()) {
				const char* sp = strcasestr(n.c_str() + std::min(prefix.size(), n.size()), suffix.c_str());
				if (!sp || strlen(sp) != suffix.size()) continue;
			}
			struct stat st;
			if (stat((prefix_dir + n).c_str(), &st) < 0) continue;
			names.push_back(prefix_dir + n);
		}
		closedir(d);
		std::sort(names.begin(), names.end());
		for (auto& n : names) { Package p; p.file = n; pk.push_back(p); printf("added: %s\n", n.c_str()); }
		printf("\n");
	}
	const int n = (int)pk.size();
	if (n == 0) { fprintf(stderr, "Error: no image to process\n"); return 1; }
	for (int i = 0; i < n; ++i) pk[i].idx = i;

	// load order (CLoadImages.cpp:71-107): zero tilt first for a tilt series
	std::vector<int> order;
	if (tomo) {
		int z = find_zero_tilt(pk);
		if (z < 0) z = n / 2;
		order.push_back(z);
		for (int i = 0; i < n; ++i) if (i != z) order.push_back(i);
	} else {
		for (int i = 0; i < n; ++i) order.push_back(i);
	}

	std::unique_ptr<SpectrumWriter> writer;
	if (!o.out_mrc.empty()) writer = std::make_unique<SpectrumWriter>();
	// full spectra come back from the GPU through page-locked memory kept for
	// the run
	struct PinnedBuf {
		float* p = nullptr;
		size_t cap = 0;
		~PinnedBuf() { if (p) gctf_host_free(p); }
	} fs_host;
	ImageLoader loader(pk, order, 4, 4);
	CRescaleImage rescale;
	CFindCtf2D find;
	int ref = -1;
	int done = 0;
	auto t_first = std::chrono::steady_clock::now();
	for (size_t k = 0; k < order.size(); ++k) {
		const int ii = order[k];
		Package& p = pk[ii];
		ImageLoader::Item im = loader.take(k);
		if (!im.ok) {
			fprintf(stderr, "Error: cannot read %s (section %d, mode %d)\n", p.file.c_str(), p.section, im.mode);
			return 1;
		}
		auto t0 = std::chrono::steady_clock::now();
		int size[2] = {im.nx, im.ny};
		try {
			// CFindSeriesCtfs::mRescaleImage + mSetupFindCtf (:108-131)
			rescale.Setup(size, o.pix);
			rescale.DoIt(im.px);
			// the next micrograph's transfer to the GPU, if it has been read,
			// overlaps this one's estimation
			ImageLoader::Item next;
			if (k + 1 < order.size() && loader.peek(k + 1, &next) && next.ok) {
				int nsize[2] = {next.nx, next.ny};
				rescale.Prefetch(next.px, nsize);
			}
			CCTFTheory theory;
			theory.Setup(o.kv, o.cs, o.ac, rescale.m_fPixSizeN, 100.0f, o.ext_phase[0] * 0.017453f);
			find.Setup1(&theory);
			find.Setup2(rescale.m_aiNewSize);
			find.GenHalfSpectrum(rescale.GetScaledImg());
		} catch (const std::exception& e) {
			fprintf(stderr, "Error: %s\n", e.what());
			return 1;
		}
		auto full = [&]() {   // mProcessFull (:155-166)
			find.SetPhase(o.ext_phase[0], fmaxf(o.ext_phase[1], 0.0f));
			find.Do2D();
			ref = ii;
		};
		auto take = [&]() {
			p.df_min = find.m_fDfMin; p.df_max = find.m_fDfMax; p.azimuth = find.m_fAstAng;
			p.ext_phase = find.m_fExtPhase; p.score = find.m_fScore;
		};
		if (done == 0) {
			full();
			take();
		} else {
			// mProcessRefine (:168-203)
			const Package& rp = pk[ref];
			float dfr[2], ast[2] = {0, 0}, ang[2] = {0, 0}, ph[2] = {0, 0};
			float range = 5000.0f * o.pix * o.pix;
			dfr[0] = 0.5f * (rp.df_min + rp.df_max);
			dfr[1] = fmaxf(dfr[0] * 0.5f, range);
			ast[0] = CFindCtfHelp::CalcAstRatio(rp.df_min, rp.df_max);
			ang[0] = rp.azimuth;
			ph[0] = rp.ext_phase;
			if (!tomo) {
				ast[1] = 0.04f;
				ang[1] = 30.0f;
				if (o.ext_phase[1] > 0) ph[1] = 40.0f;
			}
			find.Refine(dfr, ast, ang, ph);
			take();
			if (pk[ref].score < p.score) ref = ii;
			if (p.score <= 0) { full(); take(); }
		}
		// the image's copy to the GPU is done: its buffer can take the next file
		if (gctf_synchronize(SharedContext()) != GCTF_OK) {
			fprintf(stderr, "Error: %s\n", gctf_error_string(SharedContext()));
			return 1;
		}
		loader.give_back(k, im);
		// mDisplay (:215-239)
		printf("%s: %5d prcoessed, %5d left\n   Index  dfmin     dfmax    azimuth  phase   score\n"
		       "   %4d  %8.2f  %8.2f  %6.2f %6.2f %9.5f\n\n",
			base_name(p.file).c_str(), done + 1, n - 1 - done, p.idx + 1, p.df_min, p.df_max, p.azimuth,
			p.ext_phase, p.score);
		if (writer) {
			int ss[2];
			find.GetSpectSize(ss, false);
			const size_t ns = (size_t)ss[0] * ss[1];
			// a full-rate copy from the GPU, then into a recycled buffer of
			// the writer
			if (fs_host.cap < ns) {
				if (fs_host.p) gctf_host_free(fs_host.p);
				void* hp = nullptr;
				fs_host.p = gctf_host_alloc(ns * sizeof(float), &hp) == GCTF_OK ? static_cast<float*>(hp) : nullptr;
				fs_host.cap = fs_host.p ? ns : 0;
			}
			std::vector<float> fsv;
			float* fs = fs_host.p;
			if (!fs) { fsv.resize(ns); fs = fsv.data(); }
			if (find.GenFullSpectrum(fs)) {
				SpectrumWriter::Job j;
				j.nx = ss[0];
				j.ny = ss[1];
				j.px = writer->buffer(ns);
				std::memcpy(j.px.data(), fs, ns * sizeof(float));
				if (tomo || o.serial == 0) {
					j.stack = true;
					j.file = o.out_mrc;
					j.section = p.idx;
					j.nz = n;
				} else {
					std::string b = base_name(p.file), serial = b.substr(std::min(prefix.size(), b.size()));
					size_t sf = suffix.empty() ? serial.find(".mrc") : serial.find(suffix);
					if (sf != std::string::npos) serial = serial.substr(0, sf);
					j.file = serial_out(o.out_mrc, serial);
				}
				writer->push(std::move(j));
			}
		}
		double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
		printf("Processing time: %.2f (s)\n", sec);
		done += 1;
		if (done == 1) t_first = std::chrono::steady_clock::now();
		if (done % 21 == 0) { save_ctf(o, pk); save_imod(o, pk, tomo); }
	}
	if (writer) writer->finish();
	save_ctf(o, pk);
	save_imod(o, pk, tomo);
	if (done > 1) {
		// the micrographs after the first, end to end (reads, GPU, spectra
		// and tables written), without program start-up and the first
		// micrograph's plan set-up
		const double st = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_first).count();
		printf("Steady state: %d micrographs in %.4f (s) after the first\n", done - 1, st);
	}
	gctf_destroy(SharedContext());
	double tot = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
	printf("Total time:  %.3f (s)\n\n", tot);
	return 0;
}
