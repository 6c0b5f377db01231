// Device helpers shared by the gfx950 kernels: scalar traits for the four
// compute types, RNE rounding to binary16 / bfloat16, storage conversions and
// the alpha/beta/op_C/op_D epilogue.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../tk.h"

namespace tk {

// TAPP storage types (datatype.h:17-42 of the reference API).
enum { S_F32 = 0, S_F64 = 1, S_C32 = 2, S_C64 = 3, S_F16 = 4, S_BF16 = 5 };

// Compute value: real part and, for complex compute types, imaginary part.
template <typename R, bool CPLX> struct Val {
    R re, im;
};

__device__ __forceinline__ float round_bf16(float x)
{
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) {
        if (u & 0x007fffffu) u |= 0x00400000u;
        return __uint_as_float(u & 0xffff0000u);
    }
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
}

// v_cvt_f16_f32 rounds to nearest even in the default mode.
__device__ __forceinline__ float round_f16(float x) { return (float)(_Float16)x; }

template <typename R> __device__ __forceinline__ R round_mode(R x, int mode)
{
    if (mode == TK_MODE_BF16 || mode == TK_MODE_SPLIT_BF16 || mode == TK_MODE_SPLIT6_BF16)
        return (R)round_bf16((float)x);
    if (mode == TK_MODE_F16) return (R)round_f16((float)x);
    return x;
}

// Round-to-nearest-even of a binary64 value to binary16 (TK_MODE_F16) or
// bfloat16 (TK_MODE_BF16) in one step (no intermediate binary32 rounding),
// returned as the float holding it.
__device__ __forceinline__ float round16_d(double x, int mode)
{
    if (!(fabs(x) <= 1.7976931348623157e308) || x == 0.0) return (float)x;   // NaN, inf, zero
    const int p = mode == TK_MODE_F16 ? 11 : 8, emin = mode == TK_MODE_F16 ? -14 : -126;
    const double ovf = mode == TK_MODE_F16 ? 65520.0 : 3.3961775292304611e38;   // (2 - 2^-8) 2^127
    int e;
    frexp(fabs(x), &e);
    int q = e - p;
    if (q < emin - p + 1) q = emin - p + 1;
    double r = ldexp(rint(ldexp(fabs(x), -q)), q);
    if (r >= ovf) r = __builtin_inf();
    return (float)copysign(r, x);
}

__device__ __forceinline__ float bf16_bits_to_float(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ uint16_t float_to_bf16_bits(float x) { return (uint16_t)(__float_as_uint(round_bf16(x)) >> 16); }

// Reads element `off` of storage type `st` as (re, im) in double.
__device__ __forceinline__ void load_storage(const void* base, int st, int64_t off, double& re, double& im)
{
    switch (st) {
    case S_F32: re = ((const float*)base)[off]; im = 0; break;
    case S_F64: re = ((const double*)base)[off]; im = 0; break;
    case S_C32: { float2 v = ((const float2*)base)[off]; re = v.x; im = v.y; } break;
    case S_C64: { double2 v = ((const double2*)base)[off]; re = v.x; im = v.y; } break;
    case S_F16: re = (float)((const _Float16*)base)[off]; im = 0; break;
    case S_BF16: re = bf16_bits_to_float(((const uint16_t*)base)[off]); im = 0; break;
    default: re = 0; im = 0; break;
    }
}

__device__ __forceinline__ void store_storage(void* base, int st, int64_t off, double re, double im)
{
    switch (st) {
    case S_F32: ((float*)base)[off] = (float)re; break;
    case S_F64: ((double*)base)[off] = re; break;
    case S_C32: ((float2*)base)[off] = make_float2((float)re, (float)im); break;
    case S_C64: ((double2*)base)[off] = make_double2(re, im); break;
    case S_F16: ((_Float16*)base)[off] = (_Float16)round16_d(re, TK_MODE_F16); break;
    case S_BF16: ((uint16_t*)base)[off] = (uint16_t)(__float_as_uint(round16_d(re, TK_MODE_BF16)) >> 16); break;
    default: break;
    }
}

// Compute-type element access (storage == compute type).
template <typename R, bool CPLX> struct Elem;
template <typename R> struct Elem<R, false> {
    static __device__ __forceinline__ Val<R, false> ld(const void* p, int64_t off)
    {
        return {((const R*)p)[off], (R)0};
    }
    static __device__ __forceinline__ void st(void* p, int64_t off, Val<R, false> v) { ((R*)p)[off] = v.re; }
};
template <> struct Elem<float, true> {
    static __device__ __forceinline__ Val<float, true> ld(const void* p, int64_t off)
    {
        float2 v = ((const float2*)p)[off];
        return {v.x, v.y};
    }
    static __device__ __forceinline__ void st(void* p, int64_t off, Val<float, true> v)
    {
        ((float2*)p)[off] = make_float2(v.re, v.im);
    }
};
template <> struct Elem<double, true> {
    static __device__ __forceinline__ Val<double, true> ld(const void* p, int64_t off)
    {
        double2 v = ((const double2*)p)[off];
        return {v.x, v.y};
    }
    static __device__ __forceinline__ void st(void* p, int64_t off, Val<double, true> v)
    {
        ((double2*)p)[off] = make_double2(v.re, v.im);
    }
};

// op_C(C at coff), the value the epilogue scales by beta (zero when beta is
// zero).  Kernels read all the C values of a lane before any of its D
// stores: C and D may alias, so a load issued after a store has to wait for
// it.
template <typename R, bool CPLX>
__device__ __forceinline__ Val<R, CPLX> epilogue_c(const tk_gett& g, int64_t coff)
{
    Val<R, CPLX> c{(R)0, (R)0};
    if (!g.beta_zero) {
        c = Elem<R, CPLX>::ld(g.C, coff);
        if (g.round_C) {
            c.re = round_mode(c.re, g.mode);
            if (CPLX) c.im = round_mode(c.im, g.mode);
        }
        if (CPLX && g.conjC) c.im = -c.im;
    }
    return c;
}

// out = op_D(alpha * acc + beta * c) written to D at doff, c = op_C(C).
template <typename R, bool CPLX>
__device__ __forceinline__ void epilogue_d(const tk_gett& g, int64_t doff, R acc_re, R acc_im, Val<R, CPLX> c)
{
    const R ar = (R)g.alpha[0], ai = (R)g.alpha[1];
    R o_re, o_im = 0;
    if (CPLX) {
        o_re = ar * acc_re - ai * acc_im;
        o_im = ar * acc_im + ai * acc_re;
    } else {
        o_re = ar * acc_re;
    }
    if (!g.beta_zero) {
        const R br = (R)g.beta[0], bi = (R)g.beta[1];
        if (CPLX) {
            o_re += br * c.re - bi * c.im;
            o_im += br * c.im + bi * c.re;
        } else {
            o_re += br * c.re;
        }
    }
    if (CPLX && g.conjD) {
        if (g.conjD_via_f32) { o_re = (R)(float)o_re; o_im = (R)(float)o_im; }
        o_im = -o_im;
    }
    Elem<R, CPLX>::st(g.D, doff, Val<R, CPLX>{o_re, o_im});
}

// out = op_D(alpha * acc + beta * op_C(C at coff)) written to D at doff.
template <typename R, bool CPLX>
__device__ __forceinline__ void epilogue_at(const tk_gett& g, int64_t coff, int64_t doff, R acc_re, R acc_im)
{
    epilogue_d<R, CPLX>(g, doff, acc_re, acc_im, epilogue_c<R, CPLX>(g, coff));
}

// out = op_D(alpha * acc + beta * op_C(C[l, m, n])) written to D[l, m, n].
template <typename R, bool CPLX>
__device__ __forceinline__ void epilogue(const tk_gett& g, int64_t l, int64_t m, int64_t n, R acc_re, R acc_im)
{
    const int64_t coff = g.beta_zero ? 0 : g.lC[l] + g.mC[m] + g.nC[n];
    epilogue_at<R, CPLX>(g, coff, g.lD[l] + g.mD[m] + g.nD[n], acc_re, acc_im);
}

// Workgroup id -> tile id so that consecutive tiles land on the same XCD
// (workgroups are dispatched round-robin over the 8 XCDs of an MI355X).
__device__ __forceinline__ int64_t xcd_tile(int64_t bid, int64_t ntiles)
{
    const int64_t q = ntiles / 8;
    if (bid >= q * 8) return bid;
    return (bid % 8) * q + bid / 8;
}

}  // namespace tk
