// mgv_internal.hpp -- host-side helpers shared by the libmgv translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/mgv.h"

namespace mgv {

// Per-thread last error message (mgv_last_error_string).
void set_error(const char *fmt, ...);
int fail(int code, const char *fmt, ...);

struct DeviceState {
    bool init = false;
    int device = -1;
    int cus = 0;
    size_t lds_per_block = 0;
    hipStream_t default_stream = nullptr;
    void *probe_scratch = nullptr;   // this device's copy of the no-alloc probe's scratch array
};

// Resolves st.probe_scratch (the device must be current); called once per
// device by mgv_init.
int probe_scratch_setup(DeviceState &st);

// State of the calling thread's current device (mgv_init must have run).
DeviceState *current_state();

inline hipStream_t stream_of(mgv_stream_t s) {
    if (s) return (hipStream_t)s;
    DeviceState *st = current_state();
    return st ? st->default_stream : nullptr;
}

inline int hip_status(hipError_t e, const char *what) {
    if (e == hipSuccess) return MGV_OK;
    return fail((int)e, "%s: %s", what, hipGetErrorString(e));
}

// Check the launch that was just issued.
inline int check_launch(const char *what) { return hip_status(hipGetLastError(), what); }

inline bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

// Persistent grid for a kernel: CUs x resident blocks per CU, computed from
// the kernel's register and LDS use (512 VGPRs per SIMD lane, 8 waves per
// SIMD at most, 4 SIMDs and 160 KiB of LDS per CU).
template <typename K>
int persistent_blocks(K kernel, int block) {
    DeviceState *st = current_state();
    hipFuncAttributes a;
    int per_cu = 1;
    if (hipFuncGetAttributes(&a, (const void *)kernel) == hipSuccess) {
        const int waves_per_block = (block + 63) / 64;
        const int regs = a.numRegs > 0 ? ((a.numRegs + 7) / 8) * 8 : 128;
        int waves_per_simd = 512 / regs;
        if (waves_per_simd > 8) waves_per_simd = 8;
        int by_regs = waves_per_simd * 4 / waves_per_block;
        int by_lds = a.sharedSizeBytes > 0 ? (int)((160 * 1024) / a.sharedSizeBytes) : 64;
        per_cu = by_regs < by_lds ? by_regs : by_lds;
        if (per_cu < 1) per_cu = 1;
    }
    int cus = (st && st->cus > 0) ? st->cus : 256;
    return cus * per_cu;
}

}  // namespace mgv
