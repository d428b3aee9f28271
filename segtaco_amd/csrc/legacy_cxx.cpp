// legacy_cxx.cpp — C++-linkage aliases of the operator surface.
//
// The reference declares its entry points without extern "C"
// (gpu_kernels.cuh:11-93), so objects compiled against that header reference
// C++-mangled symbols (e.g. _Z10spmm_rb_srP13taco_tensor_tS0_S0_). This file
// defines those symbols by forwarding to the C-ABI dispatcher, so such
// objects link against libsegtaco unchanged. It deliberately does not include
// spmm.h, whose C-linkage declarations of the same names would conflict.
#include <stdint.h>

#include "legacy_table.h"

struct taco_tensor_t;

extern "C" void segtaco_legacy_call(int32_t id, taco_tensor_t *C,
                                    taco_tensor_t *A, taco_tensor_t *B);
extern "C" float get_gpu_timer_result(void);

// float get_gpu_timer_result() with C++ linkage (gpu_kernels.cuh:4), named
// through an assembler label because this file also sees the C declaration.
float get_gpu_timer_result_cxx() __asm__("_Z20get_gpu_timer_resultv");
float get_gpu_timer_result_cxx() { return get_gpu_timer_result(); }

#define SEGTACO_LEGACY_CXX(name, s, v, q, g, t)                     \
  void name(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B) { \
    segtaco_legacy_call(SEGTACO_LEGACY_##name, C, A, B);            \
  }
SEGTACO_LEGACY_TABLE(SEGTACO_LEGACY_CXX)
#undef SEGTACO_LEGACY_CXX
