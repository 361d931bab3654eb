// ss_tuning.h -- launch tuning of the pixel kernels, per pixel precision.
//
// *_WAVES: minimum waves per SIMD that the kernel's register allocation must
// allow (more waves hide more gather latency; fewer registers limit how many
// gathers a wave keeps in flight).  *_SAMPLES: template samples per unrolled
// chunk of a thread (a power of two).  The defaults are the values measured
// fastest on MI355X for 512x512 frames (bench.py's headline workload); any
// build may override them with -D.
#pragma once

#ifndef SS_ASSEMBLE_F32_WAVES
#define SS_ASSEMBLE_F32_WAVES 4
#endif
#ifndef SS_ASSEMBLE_F32_SAMPLES
#define SS_ASSEMBLE_F32_SAMPLES 8
#endif
#ifndef SS_ASSEMBLE_F64_WAVES
#define SS_ASSEMBLE_F64_WAVES 2
#endif
#ifndef SS_ASSEMBLE_F64_SAMPLES
#define SS_ASSEMBLE_F64_SAMPLES 4
#endif
#ifndef SS_LS_F32_WAVES
#define SS_LS_F32_WAVES 4
#endif
#ifndef SS_LS_F32_SAMPLES
#define SS_LS_F32_SAMPLES 8
#endif
#ifndef SS_LS_F64_WAVES
#define SS_LS_F64_WAVES 4
#endif
#ifndef SS_LS_F64_SAMPLES
#define SS_LS_F64_SAMPLES 4
#endif
#ifndef SS_CORR_F32_WAVES
#define SS_CORR_F32_WAVES 4
#endif
#ifndef SS_CORR_F32_SAMPLES
#define SS_CORR_F32_SAMPLES 16
#endif
#ifndef SS_CORR_F64_WAVES
#define SS_CORR_F64_WAVES 2
#endif
#ifndef SS_CORR_F64_SAMPLES
#define SS_CORR_F64_SAMPLES 8
#endif
#ifndef SS_FINAL_F32_WAVES
#define SS_FINAL_F32_WAVES 4
#endif
#ifndef SS_FINAL_F32_SAMPLES
#define SS_FINAL_F32_SAMPLES 8
#endif
#ifndef SS_FINAL_F64_WAVES
#define SS_FINAL_F64_WAVES 4
#endif
#ifndef SS_FINAL_F64_SAMPLES
#define SS_FINAL_F64_SAMPLES 4
#endif

// per-kernel values for pixel type R
#define SS_TUNE(KERNEL, R, WHAT) (sizeof(R) == 8 ? SS_##KERNEL##_F64_##WHAT : SS_##KERNEL##_F32_##WHAT)
