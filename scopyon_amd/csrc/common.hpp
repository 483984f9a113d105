// Shared device helpers for the scopyon HIP kernels (gfx950, wave64).
//
// Compiled with -ffp-contract=off: every a*b+c below rounds twice, exactly
// like the float64 numpy expressions it mirrors.  Where a fused multiply-add
// is wanted it is written as fma().
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "i1e_coeffs.hpp"

#define SMX_HD __host__ __device__ inline

// ---------------------------------------------------------------------------
// Status codes shared with include/scopyon_hip.h
// ---------------------------------------------------------------------------
#define SMX_OK 0
#define SMX_EVALUE (-1)
#define SMX_ETYPE (-2)
#define SMX_ERUNTIME (-3)
#define SMX_ENUMERIC (-4)

namespace smx {

void set_error(const char* fmt, ...);
int check_hip(hipError_t e, const char* what);

#define SMX_CHECK_LAUNCH(what)                                          \
    do {                                                                \
        int _rc = ::smx::check_hip(hipGetLastError(), what);            \
        if (_rc) return _rc;                                            \
    } while (0)

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11 "Parallel random numbers:
// as easy as 1, 2, 3").  Counter-based: (counter, key) -> 4 x 32 random bits.
// ---------------------------------------------------------------------------
struct u32x4 { uint32_t x, y, z, w; };

SMX_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)M0 * c.x;
        uint64_t p1 = (uint64_t)M1 * c.z;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = u32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// The same rounds on the device with "hi ^ word ^ key" as one v_bitop3_b32
// (three-input bitwise op, truth table 0x96 = XOR of the three; gfx950 has
// no v_xor3_b32, and the compiler otherwise emits two XORs per word).
// Results are identical to philox4x32_10.
__device__ inline u32x4 philox4x32_10_dev(u32x4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)M0 * c.x;
        uint64_t p1 = (uint64_t)M1 * c.z;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = u32x4{(uint32_t)__builtin_amdgcn_bitop3_b32(hi1, c.y, k0, 0x96), lo1,
                  (uint32_t)__builtin_amdgcn_bitop3_b32(hi0, c.w, k1, 0x96), lo0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// 53-bit uniform in [0, 1) from two 32-bit words.
SMX_HD double u01_53(uint32_t a, uint32_t b) {
    uint64_t v = ((uint64_t)a << 21) ^ (uint64_t)(b >> 11);
    return (double)(v & ((1ull << 53) - 1)) * 0x1.0p-53;
}
// 53-bit uniform in (0, 1].
SMX_HD double u01_53_open0(uint32_t a, uint32_t b) { return 1.0 - u01_53(a, b); }

// A small per-thread stream: counter (id, seq, ctr2, draw index), fixed key.
struct PhiloxStream {
    uint32_t k0, k1, id, seq, lane2, n;
    u32x4 buf;
    int left;
    SMX_HD PhiloxStream(uint64_t seed, uint32_t id_, uint32_t seq_, uint32_t lane2_)
        : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)), id(id_), seq(seq_), lane2(lane2_),
          n(0), buf{0, 0, 0, 0}, left(0) {}
    SMX_HD void refill() {
        buf = philox4x32_10(u32x4{id, seq, lane2, n}, k0, k1);
        ++n;
        left = 4;
    }
    SMX_HD uint32_t next32() {
        if (left == 0) refill();
        --left;
        return left == 3 ? buf.x : left == 2 ? buf.y : left == 1 ? buf.z : buf.w;
    }
    SMX_HD double uniform() { uint32_t a = next32(); uint32_t b = next32(); return u01_53(a, b); }
    SMX_HD double uniform_open0() { return 1.0 - uniform(); }
};

// ---------------------------------------------------------------------------
// i1e(y) = exp(-y) I1(y), y >= 0 (coefficients from gen_i1e.py).
// ---------------------------------------------------------------------------
SMX_HD double i1e(double y) {
    if (y <= SMX_I1E_YB) {
        double q = y * y * 0.25;
        double acc = smx_i1e_series[SMX_I1E_NSERIES - 1];
        for (int k = SMX_I1E_NSERIES - 2; k >= 0; --k) acc = acc * q + smx_i1e_series[k];
        return (y * 0.5) * acc * exp(-y);
    }
    double t = (2.0 * SMX_I1E_YB) / y - 1.0;
    double t2 = 2.0 * t, b1 = 0.0, b2 = 0.0;
    for (int k = SMX_I1E_NCHEB - 1; k >= 1; --k) {
        double b0 = t2 * b1 - b2 + smx_i1e_cheb[k];
        b2 = b1;
        b1 = b0;
    }
    double h = t * b1 - b2 + smx_i1e_cheb[0];
    return h / sqrt(y);
}

// ---------------------------------------------------------------------------
// numpy float semantics used by the parity paths
// ---------------------------------------------------------------------------
// numpy.remainder(a, b) for float64 (floor-division remainder, sign of b).
SMX_HD double np_remainder(double a, double b) {
    double m = fmod(a, b);
    if (m != 0.0) {
        if ((b < 0.0) != (m < 0.0)) m += b;
    } else {
        m = copysign(0.0, b);
    }
    return m;
}

// Ceil of a double, returned as a 64-bit integer (numpy.ceil(..).astype(int)).
SMX_HD long long ceil_i64(double x) { return (long long)ceil(x); }

// ceil / floor of the correctly rounded float64 quotient num / den, from the
// product q = num * inv with inv = 1 / den rounded.  q lies within 3.4e-16 |q|
// of that quotient, so the two have the same ceiling and floor unless q is
// within 1e-12 (relative) of an integer; there the division decides.
SMX_HD long long ceil_div(double num, double den, double inv) {
    const double q = num * inv;
    if (fabs(q - rint(q)) > 1e-12 * fmax(1.0, fabs(q))) return (long long)ceil(q);
    return (long long)ceil(num / den);
}
SMX_HD long long floor_div(double num, double den, double inv) {
    const double q = num * inv;
    if (fabs(q - rint(q)) > 1e-12 * fmax(1.0, fabs(q))) return (long long)floor(q);
    return (long long)floor(num / den);
}

// Read through a pointer known to address global memory.  Pointers fetched
// from by-value argument structs are otherwise treated as generic (flat)
// pointers, and flat loads share the LDS wait counter, so every LDS wait
// would also drain the loads in flight.
template <typename T>
__device__ inline T gload(const T* p) {
    return *(const __attribute__((address_space(1))) T*)(p);
}
// 16-byte form (p 16-byte aligned): one global_load_dwordx4.
__device__ inline float4 gload4(const float* p) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4 v = *(const __attribute__((address_space(1))) f4*)(p);
    return make_float4(v.x, v.y, v.z, v.w);
}

// Inclusive prefix sum over the 64 lanes of a wave on DPP lane moves (VALU,
// no LDS crossbar traffic, unlike __shfl_up's ds_bpermute): shifts by 1, 2,
// 4 and 8 inside each row of 16 lanes (lanes shifted in from outside the row
// add 0), then lane 15 of rows 0 and 2 into rows 1 and 3, then lane 31 into
// rows 2 and 3.  Call with every lane of the wave active.
__device__ inline int wave_incl_scan(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);   // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);   // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);   // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);   // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);   // row_bcast:15
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);   // row_bcast:31
    return v;
}

// Value of another lane by a DPP lane move (CTRL: a gfx9 DPP control, e.g.
// quad_perm 0xB1 = lane xor 1, 0x4E = lane xor 2 inside each quad).  Every
// lane of the wave must be active.
template <int CTRL>
__device__ inline float dpp_f32(float v) {
    return __builtin_bit_cast(
            float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, false));
}

// Summed-area-table index for a (side x side) corner grid stored either
// row-major (period == 1) or phase-major with the given period: corners whose
// row/column indices share residues mod `period` are contiguous, so one
// emitter's regular pixel-boundary grid reads a compact block.
struct SatLayout {
    int side;     // corner indices 0..side-1 on each axis (table side + 1)
    int period;   // >= 1
    int g;        // ceil(side / period)
    SMX_HD size_t index(int a, int b) const {
        if (period == 1) return (size_t)a * side + b;
        int pa = a % period, pb = b % period;
        return ((size_t)(pa * period + pb) * g + (a / period)) * g + (b / period);
    }
    // index(a, b) == row_part(a) + col_part(b)
    SMX_HD size_t row_part(int a) const {
        if (period == 1) return (size_t)a * side;
        return (size_t)(a % period) * period * g * g + (size_t)(a / period) * g;
    }
    SMX_HD size_t col_part(int b) const {
        if (period == 1) return (size_t)b;
        return (size_t)(b % period) * g * g + (size_t)(b / period);
    }
    SMX_HD size_t elems() const {
        return period == 1 ? (size_t)side * side : (size_t)period * period * g * g;
    }
};

SMX_HD SatLayout make_sat_layout(int side, int period) {
    SatLayout L;
    L.side = side;
    L.period = period < 1 ? 1 : period;
    L.g = (side + L.period - 1) / L.period;
    return L;
}

// Symmetry of the phase tables.  Along one axis an emitter of phase ph
// (0 <= ph < P) has cell boundaries 0, ph+1, ph+1+P, ... clamped to ntab, so
// its pixel d covers cells [edge(d), edge(d+1)) and there are
// np(ph) = 1 + ceil((ntab - ph - 1) / P) non-empty pixels.  The cartesian PSF
// table is symmetric under i -> ntab-1-i, which maps the boundary set of ph
// onto that of ph' = (ntab - 2 - ph) mod P with the pixels in reverse order:
// pixel d of ph integrates the same cells as pixel np-1-d of ph'.  The table is
// also symmetric under transposition, so one block per unordered pair of
// canonical phases (the smaller of ph, ph') holds every pixel integral.
struct PhaseFold {
    int P;      // period (pixel length in cells)
    int ntab;   // table side in cells
    int c0;     // (ntab - 2) mod P: ph + ph' == c0 (mod P)
    int n1;     // canonical phases in [0, c0]: 0 .. c0/2
    int nc;     // canonical phases in total
    int mf;     // np(0): the largest pixel count, block side
    // canonical rank of ph and whether ph is the reflected member of its pair
    SMX_HD int canon(int ph, bool& reflected) const {
        int partner = c0 - ph;
        if (partner < 0) partner += P;
        reflected = partner < ph;
        int c = reflected ? partner : ph;
        return c <= c0 ? c : n1 + (c - c0 - 1);
    }
    SMX_HD int uncanon(int rank) const { return rank < n1 ? rank : c0 + 1 + (rank - n1); }
    SMX_HD int npix(int ph) const {
        int phi = ph + 1;
        return phi >= ntab ? 1 : 1 + (ntab - phi + P - 1) / P;
    }
    // block of the unordered canonical pair (a <= b)
    SMX_HD size_t block(int a, int b) const {
        return (size_t)a * nc - (size_t)a * (a - 1) / 2 + (size_t)(b - a);
    }
    SMX_HD size_t elems() const { return (size_t)nc * (nc + 1) / 2 * mf * mf; }
};

SMX_HD PhaseFold make_phase_fold(int ntab, int P) {
    PhaseFold f;
    f.P = P;
    f.ntab = ntab;
    f.c0 = ((ntab - 2) % P + P) % P;
    f.n1 = f.c0 / 2 + 1;
    f.nc = f.n1 + ((f.c0 + P) / 2 - f.c0);
    f.mf = f.npix(0);
    return f;
}

// Workgroup -> tile map that keeps neighbouring tiles on one XCD.  The
// dispatcher hands workgroup b to XCD b % NXCD, so XCD x takes the contiguous
// range of tiles [x per + min(x, rem), (x + 1) per + min(x + 1, rem)) of the
// n tiles (per = n / NXCD, rem = n % NXCD): the emitters, records and table
// blocks that neighbouring tiles share are then fetched into one XCD's L2
// instead of several.  A bijection of [0, n) for every n.
constexpr int NXCD = 8;
__device__ inline int xcd_tile(int b, int n) {
    const int x = b % NXCD, i = b / NXCD, per = n / NXCD, rem = n % NXCD;
    return x * per + min(x, rem) + i;
}

}  // namespace smx
