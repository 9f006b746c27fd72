// CPU wave emulator's stand-in for the kernels' inline-assembly helpers
// (test infrastructure only): the scheduling hints have no effect on a CPU.
#ifndef RPS_EMU_GFX950_ASM_HPP
#define RPS_EMU_GFX950_ASM_HPP

#include <cstdint>

inline void keep_in_sgpr(uint32_t) {}

#endif
