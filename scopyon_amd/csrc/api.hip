// C-ABI plumbing: error strings, version, device info and host-side
// evaluations of the shared math helpers (used by CPU-only tests).
#include "common.hpp"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

namespace smx {

static thread_local char g_last_error[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

int check_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return SMX_OK;
    set_error("%s: %s", what, hipGetErrorString(e));
    return SMX_ERUNTIME;
}

}  // namespace smx

extern "C" {

const char* smx_last_error(void) { return smx::g_last_error; }

int smx_version(void) { return 10000; }

void smx_philox4x32_10_host(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    smx::u32x4 r = smx::philox4x32_10(smx::u32x4{ctr[0], ctr[1], ctr[2], ctr[3]}, key[0], key[1]);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = r.w;
}

double smx_i1e_host(double y) { return smx::i1e(y); }

double smx_np_remainder_host(double a, double b) { return smx::np_remainder(a, b); }

size_t smx_sat_index_host(int side, int period, int a, int b) {
    return smx::make_sat_layout(side, period).index(a, b);
}

int smx_device_count(int* count) {
    if (!count) {
        smx::set_error("smx_device_count: null pointer");
        return SMX_EVALUE;
    }
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return SMX_OK;
}

}  // extern "C"
