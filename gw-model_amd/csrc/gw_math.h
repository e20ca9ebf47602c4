// gw_math.h -- arithmetic shared by the gfx950 kernels and the host side of
// the library: libm, the exact flux accumulator, the uniform map and the
// xoshiro/xoroshiro engines. Everything here is __host__ __device__ and uses
// only IEEE-754 operations with a single rounding each (+ - * / sqrt fma,
// rint, ldexp), so host and device produce identical bits. The library is
// compiled with -ffp-contract=off; fused operations appear only as explicit
// fma() calls.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GW_HD __host__ __device__ __forceinline__

namespace gw {

GW_HD uint64_t bits_of(double d) { return static_cast<uint64_t>(__builtin_bit_cast(uint64_t, d)); }
GW_HD double double_of(uint64_t u) { return __builtin_bit_cast(double, u); }
GW_HD uint64_t rotl64(uint64_t x, int s) { return (x << s) | (x >> (64 - s)); }

// ---------------------------------------------------------------------------
// exp: Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, degree-13 Taylor
// polynomial of e^r in Horner form with fused multiply-adds, scaling by 2^k.
// ---------------------------------------------------------------------------
GW_HD double pow2i(int k) { return double_of(static_cast<uint64_t>(k + 1023) << 52); }

GW_HD double exp(double x)
{
    if (x != x) return x;
    if (x > 709.79) return __builtin_inf();
    if (x < -745.2) return 0.0;
    const double kd = __builtin_rint(x * 0x1.71547652b82fep+0);
    double r = __builtin_fma(-kd, 0x1.62e42fefa39efp-1, x);
    r = __builtin_fma(-kd, 0x1.abc9e3b39803fp-56, r);
    double p = 0x1.6124613a86d09p-33;
    p = __builtin_fma(p, r, 0x1.1eed8eff8d898p-29);
    p = __builtin_fma(p, r, 0x1.ae64567f544e4p-26);
    p = __builtin_fma(p, r, 0x1.27e4fb7789f5cp-22);
    p = __builtin_fma(p, r, 0x1.71de3a556c734p-19);
    p = __builtin_fma(p, r, 0x1.a01a01a01a01ap-16);
    p = __builtin_fma(p, r, 0x1.a01a01a01a01ap-13);
    p = __builtin_fma(p, r, 0x1.6c16c16c16c17p-10);
    p = __builtin_fma(p, r, 0x1.1111111111111p-7);
    p = __builtin_fma(p, r, 0x1.5555555555555p-5);
    p = __builtin_fma(p, r, 0x1.5555555555555p-3);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    const int k = static_cast<int>(kd);
    if (k > 1023) return (p * pow2i(1023)) * pow2i(k - 1023);
    if (k < -1021) return (p * pow2i(k + 1000)) * pow2i(-1000);
    return p * pow2i(k);
}

// ---------------------------------------------------------------------------
// log: x = 2^e m with m in [sqrt(2)/2, sqrt(2)); log m = 2 atanh(s),
// s = (m - 1)/(m + 1), atanh series through s^25 (Horner, fused).
// ---------------------------------------------------------------------------
GW_HD double log(double x)
{
    if (x != x) return x;
    if (x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -__builtin_inf();
    if (x == __builtin_inf()) return x;
    uint64_t u = bits_of(x);
    int e = static_cast<int>((u >> 52) & 0x7ff);
    if (e == 0) {
        u = bits_of(x * 0x1p54);
        e = static_cast<int>((u >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    double m = double_of((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > 0x1.6a09e667f3bcdp+0) {
        m = m * 0.5;
        e += 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double R = 0x1.47ae147ae147bp-5;
    R = __builtin_fma(R, z, 0x1.642c8590b2164p-5);
    R = __builtin_fma(R, z, 0x1.8618618618618p-5);
    R = __builtin_fma(R, z, 0x1.af286bca1af28p-5);
    R = __builtin_fma(R, z, 0x1.e1e1e1e1e1e1ep-5);
    R = __builtin_fma(R, z, 0x1.1111111111111p-4);
    R = __builtin_fma(R, z, 0x1.3b13b13b13b14p-4);
    R = __builtin_fma(R, z, 0x1.745d1745d1746p-4);
    R = __builtin_fma(R, z, 0x1.c71c71c71c71cp-4);
    R = __builtin_fma(R, z, 0x1.2492492492492p-3);
    R = __builtin_fma(R, z, 0x1.999999999999ap-3);
    R = __builtin_fma(R, z, 0x1.5555555555555p-2);
    R = R * z;
    const double s2 = 2.0 * s;
    const double lm = __builtin_fma(s2, R, s2);
    const double ed = static_cast<double>(e);
    return __builtin_fma(ed, 0x1.62e42fefa39efp-1, __builtin_fma(ed, 0x1.abc9e3b39803fp-56, lm));
}

GW_HD double pow(double x, double y)
{
    if (y == 0.0) return 1.0;
    if (x != x || y != y) return __builtin_nan("");
    if (x < 0.0) return __builtin_nan("");
    if (x == 0.0) return (y > 0.0) ? 0.0 : __builtin_inf();
    return gw::exp(y * gw::log(x));
}

// ---------------------------------------------------------------------------
// Exact flux accumulator: 128-bit two's-complement fixed point with 2^-80
// resolution. Terms are added exactly (truncated below 2^-80), so a sum does
// not depend on the order of its terms; it is rounded to double once.
// ---------------------------------------------------------------------------
struct Fx {
    uint64_t lo, hi;
};

GW_HD void fx_zero(Fx &a) { a.lo = 0; a.hi = 0; }

GW_HD void fx_add_raw(Fx &a, uint64_t lo, uint64_t hi)
{
    const uint64_t nlo = a.lo + lo;
    a.hi = a.hi + hi + (nlo < lo ? 1u : 0u);
    a.lo = nlo;
}

GW_HD void fx_add(Fx &a, const Fx &b) { fx_add_raw(a, b.lo, b.hi); }

// returns a nonzero flag for NaN/Inf (1) or |x| >= 2^24 (2)
GW_HD int fx_add(Fx &a, double x)
{
    const uint64_t u = bits_of(x);
    const int E = static_cast<int>((u >> 52) & 0x7ff);
    if (E == 0) return 0;
    if (E == 0x7ff) return 1;
    if (E >= 1047) return 2;
    const uint64_t M = (u & 0x000fffffffffffffULL) | (1ULL << 52);
    const int sh = E - 995;
    uint64_t lo, hi;
    if (sh >= 64) {
        lo = 0;
        hi = M << (sh - 64);
    } else if (sh > 0) {
        lo = M << sh;
        hi = M >> (64 - sh);
    } else if (sh == 0) {
        lo = M;
        hi = 0;
    } else {
        lo = (sh > -53) ? (M >> (-sh)) : 0;
        hi = 0;
    }
    if (u >> 63) {  // negate the 128-bit magnitude
        lo = ~lo + 1;
        hi = ~hi + (lo == 0 ? 1u : 0u);
    }
    fx_add_raw(a, lo, hi);
    return 0;
}

GW_HD int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

GW_HD double fx_to_double(const Fx &a)
{
    uint64_t lo = a.lo, hi = a.hi;
    if ((lo | hi) == 0) return 0.0;
    const bool neg = (hi >> 63) != 0;
    if (neg) {
        lo = ~lo + 1;
        hi = ~hi + (lo == 0 ? 1u : 0u);
    }
    const int p = hi ? 127 - clz64(hi) : 63 - clz64(lo);
    double d;
    if (p <= 52) {
        d = __builtin_ldexp(static_cast<double>(lo), -80);
    } else {
        int shift = p - 52;
        uint64_t m, rem_hi_bit, sticky;
        // m = mag >> shift; rem = mag & (2^shift - 1); half = 2^(shift-1)
        if (shift >= 64) {
            m = hi >> (shift - 64);
            const int s2 = shift - 64;  // bits of hi below m, then all of lo
            const uint64_t hrem = s2 ? (hi & ((1ULL << s2) - 1)) : 0;
            rem_hi_bit = s2 ? ((hrem >> (s2 - 1)) & 1) : (lo >> 63);
            sticky = s2 ? ((hrem & ((1ULL << (s2 - 1)) - 1)) | lo) : (lo & ~(1ULL << 63));
        } else {
            m = (lo >> shift) | (shift ? (hi << (64 - shift)) : 0);
            const uint64_t rem = lo & ((1ULL << shift) - 1);
            rem_hi_bit = (rem >> (shift - 1)) & 1;
            sticky = rem & ((1ULL << (shift - 1)) - 1);
        }
        if (rem_hi_bit && (sticky || (m & 1))) {
            ++m;
            if (m == (1ULL << 53)) {
                m >>= 1;
                ++shift;
            }
        }
        d = __builtin_ldexp(static_cast<double>(m), shift - 80);
    }
    return neg ? -d : d;
}

// ---------------------------------------------------------------------------
// Uniform map of std::uniform_real_distribution<double>(0,1) over a 64-bit
// engine (libstdc++ generate_canonical): round(x)/2^64, clamped below 1.
// The u64 -> f64 conversion is split so that it rounds exactly once.
// ---------------------------------------------------------------------------
GW_HD double u64_to_f64_rn(uint64_t x)
{
    const double hi = static_cast<double>(static_cast<uint32_t>(x >> 32)) * 4294967296.0;
    const double lo = static_cast<double>(static_cast<uint32_t>(x));
    return hi + lo;
}

GW_HD double u01(uint64_t x)
{
    double u = u64_to_f64_rn(x) * 0x1p-64;
    if (u >= 1.0) u = 0x1.fffffffffffffp-1;
    return u;
}

// FP32 uniform in [0, 1): top 24 bits, exact.
GW_HD float u01f(uint64_t x) { return static_cast<float>(static_cast<uint32_t>(x >> 40)) * 0x1p-24f; }

// ---------------------------------------------------------------------------
// PRNG engines (ids of include/gw_capi.h). State is always 4 x u64; the
// 128-bit engines use the first two words.
// ---------------------------------------------------------------------------
GW_HD uint64_t splitmix64_next(uint64_t &s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

GW_HD void rng_seed(int prng, uint64_t s[4], uint64_t seed)
{
    uint64_t sm = seed;
    s[0] = splitmix64_next(sm);
    s[1] = splitmix64_next(sm);
    if (prng >= 3) {
        s[2] = 0;
        s[3] = 0;
    } else {
        s[2] = splitmix64_next(sm);
        s[3] = splitmix64_next(sm);
    }
}

template <int PRNG>
GW_HD uint64_t rng_next(uint64_t &s0, uint64_t &s1, uint64_t &s2, uint64_t &s3)
{
    uint64_t result;
    if constexpr (PRNG <= 2) {
        if constexpr (PRNG == 0) result = rotl64(s1 * 5, 7) * 9;
        else if constexpr (PRNG == 1) result = rotl64(s0 + s3, 23) + s0;
        else result = s0 + s3;
        const uint64_t t = s1 << 17;
        s2 ^= s0;
        s3 ^= s1;
        s1 ^= s2;
        s0 ^= s3;
        s2 ^= t;
        s3 = rotl64(s3, 45);
    } else {
        const uint64_t a = s0;
        uint64_t b = s1;
        if constexpr (PRNG == 3) result = rotl64(a * 5, 7) * 9;
        else if constexpr (PRNG == 4) result = rotl64(a + b, 17) + a;
        else result = a + b;
        b ^= a;
        if constexpr (PRNG == 4) {
            s0 = rotl64(a, 49) ^ b ^ (b << 21);
            s1 = rotl64(b, 28);
        } else {
            s0 = rotl64(a, 24) ^ b ^ (b << 16);
            s1 = rotl64(b, 37);
        }
    }
    return result;
}

// Stream key of CRU i of a cell seeded with `seed` (the SSA stream); the
// rest-state sampler uses the complement of the key.
GW_HD uint64_t stream_key(uint64_t seed, int64_t i) { return seed + (static_cast<uint64_t>(i) << 32); }

}  // namespace gw
