// Portability layer: the kernel bodies compile for gfx950 (hipcc) and, with
// SFX_EMULATE defined, for the host CPU, where tests run one work-group as
// std::threads with a barrier in place of __syncthreads and check every
// global and LDS index against its buffer.
#pragma once
#include <stdint.h>

#ifdef SFX_EMULATE
#include <cmath>
#include <cstdio>
#include <cstdlib>
struct double2 { double x, y; };
inline double2 make_double2(double x, double y) { return double2{x, y}; }
using std::fma;
#define SFX_DEV inline
#define SFX_DEV_NOINLINE inline
#define __restrict__
void sfx_emu_sync();
void sfx_emu_wave_sync();
[[noreturn]] void sfx_emu_oob(const char* what, long long idx, long long n, const char* file, int line);
#define SFX_SYNC() sfx_emu_sync()
#define SFX_WAVE_SYNC() sfx_emu_wave_sync()
#define SFX_OPAQUE(x) do { } while (0)
#define SFX_OPAQUE_S(x) do { } while (0)
#define SFX_KARG
#define SFX_SCHED_BARRIER() do { } while (0)
#define SFX_CHECK(what, idx, n) \
  do { if ((long long)(idx) < 0 || (long long)(idx) >= (long long)(n)) sfx_emu_oob(what, (long long)(idx), (long long)(n), __FILE__, __LINE__); } while (0)
#else
#include <hip/hip_runtime.h>
#define SFX_DEV __device__ __forceinline__
// a device function compiled on its own (own register allocation and stack frame)
#define SFX_DEV_NOINLINE __device__ __noinline__
#define SFX_SYNC() __syncthreads()
// Exchange inside one wave through its private LDS region: a wave's LDS
// instructions execute in order, so only compiler reordering must be stopped.
#define SFX_WAVE_SYNC()                                   \
  do {                                                    \
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront"); \
    __builtin_amdgcn_wave_barrier();                      \
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront"); \
  } while (0)
#define SFX_CHECK(what, idx, n) do { } while (0)
// Hide a value's origin from the optimizer, so that what is computed from it
// is recomputed where it is used instead of being kept live (or spilled)
// across a long stretch of code.  Emits no instruction.
#define SFX_OPAQUE(x) __asm__ volatile("" : "+v"(x))
// The same for a wave-uniform value kept in scalar registers (a pointer
// read through redefined at this point: loads through it are not hoisted
// above it).
#define SFX_OPAQUE_S(x) __asm__ volatile("" : "+s"(x))
// Address space of the kernel arguments: uniform loads from it are scalar loads.
#define SFX_KARG __attribute__((address_space(4)))
// Keep the instruction scheduler from moving code across this point (for
// example loads hoisted above it, whose destinations would be live early).
#define SFX_SCHED_BARRIER() __builtin_amdgcn_sched_barrier(0)
#endif
