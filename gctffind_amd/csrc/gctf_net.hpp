// gctf_net.hpp -- the defocus-proposal network (no reference counterpart).
//
// A small 1D convolutional network reads the first 256 bins of a
// micrograph's radial spectrum (the background-removed, low-passed
// half spectrum averaged over rings, CFindCtf1D.cpp:113-117) and predicts
// (ln(mean defocus / 1e4 A), astigmatism / 1e3 A, azimuth in rad,
// confidence logit).  Layers:
//   input   bins [16, 256) z-scored per spectrum, bins [0, 16) set to 0
//   conv1   1 -> 64 channels, 7 taps, zero padding 3, ReLU, mean-pool 2
//   conv2-4 64 -> 64 channels, 7 taps, zero padding 3, ReLU, mean-pool 2
//   fc1     1024 (= 64 channels x 16 positions) -> 128, ReLU
//   fc2     128 -> 4
// Activations are channel-major: [channel][batch * length].  The GEMMs run
// on MFMA with BF16 operands (FP32 accumulation, FP32 master parameters);
// an FP32-operand MFMA variant serves gradient checks.
#pragma once
#include <hip/hip_runtime.h>

namespace gctf {

constexpr int kNetBins = 256;                 // radial bins read
constexpr int kNetBin0 = 16;                  // first bin of the z-scored input
constexpr int kNetCh = 64, kNetTaps = 7, kNetHidden = 128, kNetOut = 4;
constexpr int kNetL4 = kNetBins >> 4;         // length after four 2x pools
constexpr int kNetFlat = kNetCh * kNetL4;     // 1024
constexpr int kNetCol = kNetCh * kNetTaps;    // 448: im2col rows of conv2-4

// Offsets in the flat FP32 parameter vector.
constexpr long long kNetW1 = 0;
constexpr long long kNetB1 = kNetW1 + kNetCh * kNetTaps;
constexpr long long kNetW2 = kNetB1 + kNetCh;
constexpr long long kNetB2 = kNetW2 + (long long)kNetCh * kNetCol;
constexpr long long kNetW3 = kNetB2 + kNetCh;
constexpr long long kNetB3 = kNetW3 + (long long)kNetCh * kNetCol;
constexpr long long kNetW4 = kNetB3 + kNetCh;
constexpr long long kNetB4 = kNetW4 + (long long)kNetCh * kNetCol;
constexpr long long kNetW5 = kNetB4 + kNetCh;
constexpr long long kNetB5 = kNetW5 + (long long)kNetHidden * kNetFlat;
constexpr long long kNetW6 = kNetB5 + kNetHidden;
constexpr long long kNetB6 = kNetW6 + (long long)kNetOut * kNetHidden;
constexpr long long kNetParams = kNetB6 + kNetOut;

// Device activations of one batch size (forward, and backward when
// `train`): allocated once, reused by every call.
struct NetBuffers;
hipError_t net_buffers_create(int max_batch, bool train, NetBuffers** out);
void net_buffers_destroy(NetBuffers* b);
int net_buffers_capacity(const NetBuffers* b);

// Forward pass of n <= capacity spectra: radial [n][ld] floats (bins 0..255
// used); out [n][4] raw network outputs.  Launches only (no allocation, no
// synchronisation), so it can be captured in a graph.
hipError_t net_forward(NetBuffers* b, const float* params, const float* radial, int ld, int n, bool bf16,
	float* out, hipStream_t s);

}  // namespace gctf
