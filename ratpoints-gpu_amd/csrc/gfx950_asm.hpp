// Inline-assembly helpers of the gfx950 sieve kernels, kept in one header so
// the kernels themselves hold no assembly.
#ifndef RATPOINTS_GPU_AMD_GFX950_ASM_HPP
#define RATPOINTS_GPU_AMD_GFX950_ASM_HPP

#include <hip/hip_runtime.h>

#include <cstdint>

// Keeps a wave-uniform value in an SGPR at this point of the program: the
// empty asm takes it as a scalar operand, so the compiler must have computed
// it here instead of sinking the computation past later code.
__device__ __forceinline__ void keep_in_sgpr(uint32_t v) { asm volatile("" ::"s"(v)); }

#endif
