// Helpers shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sfx_kernels.h"

namespace sfx {

int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* where);
void clear_err();
int get_twiddles(int MP, Twiddles* tw);
int staging(int slot, size_t bytes, void** out);
double now_s();
int check_modes(const char* name, int64_t np, int32_t m_max_c, int32_t m_max_pad);
int check_procs(const char* name, int64_t np, int32_t nb_procs, int64_t bloc_size);

}  // namespace sfx
