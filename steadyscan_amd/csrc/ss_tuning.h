// ss_tuning.h -- launch variants of the pixel kernels, per pixel precision.
//
// Each of the assembly, line-search and correlation kernels is built in
// kVariants variants of (samples, waves):
//   samples: template samples a thread keeps in flight (the unrolled chunk
//            of rows times the samples per row; a power of two),
//   waves:   minimum waves per SIMD that the kernel's register allocation
//            must allow (more waves hide more gather latency; fewer
//            registers limit how many gathers a wave keeps in flight).
// Variant 0 is the default for batches that fill the chip: the values
// measured fastest on MI355X for batches of 512x512 groups (the headline
// workload); in FP32 the assembly and line search take 4 samples, which
// measure as fast as 8 without spilling registers.  A batch of at most 8
// block columns per CU (one 512x512 group of 50 frames has 1600) runs too
// few workgroups to fill the chip, and there kSmallBatchVariant picks the
// variants measured fastest there: FP64 assembly and correlation at 4 waves
// per SIMD (one 512x512 group's first tick 142 and 72 us against 154 and 75)
// and the FP32 line search with 8 samples in flight (131 against 141 us).
// The engine's autotuner (off unless asked for) times all variants on a
// group shape's first tick and keeps the fastest per kernel.
// A variant only changes how many rows a thread unrolls and how registers
// are allocated: every thread still accumulates its pixels in the same
// order, so all variants give the same results bit for bit.
#pragma once

namespace ss {

constexpr int kVariants = 3;
enum TuneKernel { TK_ASSEMBLE = 0, TK_LS = 1, TK_CORR = 2, TK_COUNT = 3 };

struct TuneVariant {
    int samples, waves;
};

// [kernel][fp64][variant]
constexpr TuneVariant kTuneTable[TK_COUNT][2][kVariants] = {
    // assembly: FP32, FP64
    {{{4, 4}, {8, 4}, {8, 3}}, {{4, 2}, {2, 3}, {2, 4}}},
    // line search
    {{{4, 4}, {8, 4}, {8, 3}}, {{4, 4}, {2, 4}, {8, 3}}},
    // correlation (one sample per row; a chunk holds at most 8 rows)
    {{{8, 4}, {4, 4}, {2, 4}}, {{8, 2}, {4, 3}, {2, 3}}},
};

// [kernel][fp64]: the variant of batches of at most 8 block columns per CU
constexpr int kSmallBatchVariant[TK_COUNT][2] = {{0, 2}, {1, 0}, {0, 2}};

template <int KERNEL, typename R, int V>
__host__ __device__ constexpr TuneVariant tune_variant() {
    return kTuneTable[KERNEL][sizeof(R) == 8 ? 1 : 0][V];
}

// the final pass runs once per group: one variant
constexpr int kFinalSamplesF32 = 4, kFinalWavesF32 = 4, kFinalSamplesF64 = 4, kFinalWavesF64 = 4;

}  // namespace ss
