// fft.hpp -- the batched FFT's descriptor tables, shared by fft.hip (plans)
// and serve.hip (served batches that build their tables per batch).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace aeth {
namespace fft {

struct Desc {
    uint32_t first_sig; // index into the signal table
    uint16_t count;     // signals in this chunk
    uint16_t log2n;
    uint32_t stride;    // elements between consecutive points of a signal (1 = contiguous)
};

struct Sig {
    uint64_t in_off;  // complex elements
    uint64_t out_off; // complex elements
};

// Descriptors and signal table resident on the device.
struct BatchPlan {
    int prec = 0; // 0 complex128, 1 complex64, 2 complex64 with float-float arithmetic, 3 complex64 in memory with FP64 arithmetic
    int dev = 0;
    Desc* descs = nullptr;
    Sig* sigs = nullptr;
    uint32_t ndesc = 0;
    uint64_t nsig = 0;
    uint32_t stride = 1; // of every descriptor (1: contiguous signals)
};

uint32_t log2_exact(uint64_t n);
// Group signals by length, largest first; a chunk holds up to 8192 points of one length.
void build_tables(const uint32_t* log2n, const uint64_t* in_off, const uint64_t* out_off, uint64_t count,
                  std::vector<Desc>& descs, std::vector<Sig>& sigs, uint32_t stride = 1);
// One launch over all descriptors of the plan. norm 0/1/2: numpy norms per length; 3: fixed_scale.
void launch(const BatchPlan& p, const void* in, void* out, int forward, int norm, hipStream_t st,
            double fixed_scale = 1.0);
// The same transforms (contiguous signals; prec 0 complex128 or 2 complex64 float-float) with a fixed
// persistent grid that reads the descriptor count from *ndesc on the device: capturable once, valid for
// any table. persistent_grid(prec) is the grid that fills the device; prepare() builds the twiddle tables
// (call both before a capture).
void prepare(hipStream_t st);
int persistent_grid(int prec);
void launch_persistent(int prec, int grid, const Desc* descs, const Sig* sigs, const uint32_t* ndesc, const void* in,
                       void* out, int forward, int norm, hipStream_t st);

} // namespace fft
} // namespace aeth
