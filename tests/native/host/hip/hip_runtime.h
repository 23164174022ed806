// Host (CPU) build of the library's field and curve headers for the native
// CPU tests: the device qualifiers become plain inline functions.
#pragma once
#define __device__
#define __forceinline__ inline __attribute__((always_inline))
#define __noinline__ __attribute__((noinline))
#define __constant__
#define __host__
