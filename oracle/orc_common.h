/*
 * ORACLE -- test infrastructure only.
 *
 * CPU restatement of the APPFIS reference OpenMP runtime (Appfis/Runtime.hpp,
 * Appfis/Grid.hpp, Appfis/Communication.hpp of the reference) driving the config
 * stencils.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and only as the checker or the timed CPU baseline; the product path
 * (libappfis_hip.so, Appfis/ headers) never links or calls it.
 *
 * Shared definitions for orc_*.c.
 */
#ifndef ORC_COMMON_H
#define ORC_COMMON_H

#include <stdint.h>
#include <stddef.h>

/* element types: identical numbering to include/appfis_hip.h */
enum { ORC_U8 = 0, ORC_I32 = 1, ORC_F32 = 2, ORC_F64 = 3, ORC_F16 = 4, ORC_BF16 = 5, ORC_I64 = 6 };
/* APPFIS::REDUCTION (reference Appfis/Env.hpp:32) */
enum { ORC_MAX = 0, ORC_MIN = 1, ORC_SUM = 2 };
/* APPFIS::DIRECTION (reference Appfis/Env.hpp:30) */
enum { ORC_LEFT = 0, ORC_RIGHT = 1, ORC_UP = 2, ORC_DOWN = 3, ORC_FRONT = 4, ORC_BACK = 5 };

static inline size_t orc_esize(int dtype)
{
    switch (dtype) {
    case ORC_U8: return 1;
    case ORC_I32: return 4;
    case ORC_F32: return 4;
    case ORC_F64: return 8;
    case ORC_F16: return 2;
    case ORC_BF16: return 2;
    case ORC_I64: return 8;
    default: return 0;
    }
}

/* splitmix64 finaliser (Steele, Lea, Flood: "Fast splittable pseudorandom number
 * generators", OOPSLA 2014); r(seed, n) is the n-th output of the sequence seeded
 * with `seed`. */
static inline uint64_t orc_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static inline uint64_t orc_rng(uint64_t seed, uint64_t n)
{
    return orc_mix64(seed + (n + 1) * 0x9E3779B97F4A7C15ULL);
}
static inline double orc_u53(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }
static inline float orc_u24(uint64_t r) { return (float)(r >> 40) * (1.0f / 16777216.0f); }

/* IEEE binary16 / bfloat16 <-> binary32, round to nearest even (portable C). */
static inline uint32_t orc_f2u(float f) { union { float f; uint32_t u; } c; c.f = f; return c.u; }
static inline float orc_u2f(uint32_t u) { union { float f; uint32_t u; } c; c.u = u; return c.f; }

static inline uint16_t orc_f32_to_bf16(float f)
{
    uint32_t u = orc_f2u(f);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
static inline float orc_bf16_to_f32(uint16_t h) { return orc_u2f((uint32_t)h << 16); }

static inline uint16_t orc_f32_to_f16(float f)
{
    uint32_t u = orc_f2u(f);
    uint32_t sign = (u >> 16) & 0x8000u;
    uint32_t absu = u & 0x7FFFFFFFu;
    if (absu >= 0x7F800000u) /* inf or nan */
        return (uint16_t)(sign | 0x7C00u | (absu > 0x7F800000u ? 0x200u : 0u));
    if (absu >= 0x477FF000u) /* rounds to >= 65520 -> inf */
        return (uint16_t)(sign | 0x7C00u);
    if (absu < 0x38800000u) { /* result is subnormal or zero in binary16 */
        /* value = absu as float; subnormal step 2^-24 */
        float a = orc_u2f(absu);
        /* a * 2^24 rounded to nearest even integer */
        float scaled = a * 16777216.0f; /* exact: power-of-two scaling, < 2^10 */
        uint32_t q = (uint32_t)scaled;
        float frac = scaled - (float)q;
        if (frac > 0.5f || (frac == 0.5f && (q & 1u))) q++;
        return (uint16_t)(sign | q);
    }
    /* normal range: rebias exponent, round mantissa 23 -> 10 bits */
    uint32_t mant = absu & 0x007FFFFFu;
    int32_t exp = (int32_t)(absu >> 23) - 127 + 15;
    uint32_t hm = mant >> 13;
    uint32_t rem = mant & 0x1FFFu;
    uint32_t out = ((uint32_t)exp << 10) | hm;
    if (rem > 0x1000u || (rem == 0x1000u && (hm & 1u))) out++; /* carry may bump exponent: still correct */
    return (uint16_t)(sign | out);
}
static inline float orc_f16_to_f32(uint16_t h)
{
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t exp = ((uint32_t)h >> 10) & 0x1Fu;
    uint32_t mant = (uint32_t)h & 0x3FFu;
    if (exp == 0) {
        if (mant == 0) return orc_u2f(sign);
        float v = (float)mant * (1.0f / 16777216.0f); /* mant * 2^-24, exact */
        return sign ? -v : v;
    }
    if (exp == 31) return orc_u2f(sign | 0x7F800000u | (mant << 13));
    return orc_u2f(sign | ((exp - 15 + 127) << 23) | (mant << 13));
}

#endif
