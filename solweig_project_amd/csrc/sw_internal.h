// Internal helpers shared by the kernel and runtime translation units of
// libsolweig_hip.so (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/solweig_hip.h"

#define SW_HIDDEN __attribute__((visibility("hidden")))

extern "C" {
// record a message for sw_last_error() and return code
SW_HIDDEN int swi_set_err(int code, const char *msg);
// swi_set_err with fmt's one %s replaced by who
SW_HIDDEN int swi_fmt_err(int code, const char *fmt, const char *who);
// SW_E_HIP with "<where>: <hip error>" unless e == hipSuccess
SW_HIDDEN int swi_hip_check(hipError_t e, const char *where);
// SW_E_ARG unless 0 < rows, cols and rows * cols < 2^31
SW_HIDDEN int swi_check_shape(int rows, int cols, const char *who);
// checks of a batch's host items (shapes, pointers, sw_batch_layout's block
// ranges); returns the grid totals of the sun-march and flux phases
SW_HIDDEN int swi_check_batch(const sw_batch_item_t *items, int nitems, long *sun_total, long *flux_total,
                              const char *who);
}

static inline hipStream_t swi_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }
