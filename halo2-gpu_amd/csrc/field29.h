// BN254 field elements in registers as 9 limbs of 29 bits, for the kernels
// whose time is spent in field multiplications (NTT passes, MSM buckets).
//
// Why 29 bits: a 29x29-bit product is < 2^58, so a column of up to 18
// partial products (9 of a*b, 9 of m*p) plus the incoming carry stays below
// 2^64 and accumulates with one `v_mad_u64_u32` each and no carry handling.
// The 32-bit product-scanning form needs a carry capture after every
// multiply-add; this one issues ~230 instructions per product instead of
// ~370 (plus the hazard waits of the SGPR carry), measured 172 vs 110-123
// G products/s on gfx950.
//
// Montgomery radix R' = 2^261 (9 x 29 bits): mul(a, b) = a * b / 2^261 mod m.
// Storage stays the boundary format (32 B, Montgomery a * 2^256, reduced);
// kernels keep data in the 2^256 scale and hold their constants (twiddles,
// coset powers, scale factors) in the 2^261 scale, so that
// mul(x * 2^256, w * 2^261) = x * w * 2^256: no per-element conversion.
//
// Reduction is lazy. mul() accepts one operand with limbs < 2^31 and the
// other < 2^29 (or both < 2^30) and any values with a * b < m * 2^261, and
// returns normalised limbs (each < 2^29) of a value < m + a * b / 2^261.
// add() and sub2() work limb by limb without carries; callers bound values
// and limb sizes (see the NTT pass kernel) and reduce_full() brings a
// normalised value < 2^261 to the canonical [0, m).
#pragma once

#include "field.h"

namespace h2 {

struct Fr29Params {
    static constexpr uint32_t M[9] = { 0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u,
        0x00a6e141u, 0x0e5c2634u, 0x0030644eu };
    static constexpr uint32_t PINV = 0x0fffffffu; // -m^-1 mod 2^29
    // 2m with every limb but the top >= 2^29 - 1 (a + K2M - b never borrows)
    static constexpr uint32_t K2M[9] = { 0x20000002u, 0x3e1f593eu, 0x3cb848a0u, 0x2fa121e5u, 0x2b0ba505u,
        0x25b68180u, 0x214dc281u, 0x3cb84c67u, 0x0060c89bu };
    static constexpr uint32_t K4M[9] = { 0x20000004u, 0x3c3eb27du, 0x39709142u, 0x3f4243ccu, 0x36174a0bu,
        0x2b6d0301u, 0x229b8503u, 0x397098cfu, 0x00c19138u };
    // 2^261 mod m (the 2^261-scale one)
    static constexpr uint32_t ONE[9] = { 0x0fffff57u, 0x1ea70ab4u, 0x052c068bu, 0x17504f49u, 0x0aa8075bu,
        0x1d4240ceu, 0x11d54c07u, 0x052ac7a8u, 0x000dc836u };
    // 2^261 mod m as a 256-bit integer: Fp256::mul(x * 2^256, TO261) = x * 2^261
    static constexpr uint32_t TO261[8] = { 0x8fffff57u, 0x2fd4e156u, 0xa494b01au, 0x75bba827u, 0x819caa80u,
        0x5301fa84u, 0x563d4475u, 0x0dc83629u };
    // 2^783 mod m in 29-bit limbs: mul(A^-1, C783) = A^-1 * 2^522
    static constexpr uint32_t C783[9] = { 0x001fddb2u, 0x17d30b63u, 0x1a2600eeu, 0x09507c47u, 0x1496b29bu,
        0x0b00a268u, 0x15b645ebu, 0x1f9fcb3du, 0x001baa96u };
};

struct Fq29Params {
    static constexpr uint32_t M[9] = { 0x187cfd47u, 0x010460b6u, 0x1c72a34fu, 0x02d522d0u, 0x1585d978u, 0x02db40c0u,
        0x00a6e141u, 0x0e5c2634u, 0x0030644eu };
    static constexpr uint32_t PINV = 0x04866389u;
    static constexpr uint32_t K2M[9] = { 0x30f9fa8eu, 0x2208c16cu, 0x38e5469du, 0x25aa45a0u, 0x2b0bb2efu,
        0x25b68180u, 0x214dc281u, 0x3cb84c67u, 0x0060c89bu };
    static constexpr uint32_t K4M[9] = { 0x21f3f51cu, 0x241182dau, 0x31ca8d3bu, 0x2b548b42u, 0x361765dfu,
        0x2b6d0301u, 0x229b8503u, 0x397098cfu, 0x00c19138u };
    static constexpr uint32_t ONE[9] = { 0x157ccc21u, 0x141c2758u, 0x185230d3u, 0x014c0419u, 0x0aa36fb9u,
        0x1d4240ceu, 0x11d54c07u, 0x052ac7a8u, 0x000dc836u };
    static constexpr uint32_t TO261[8] = { 0x157ccc21u, 0x4e8384ebu, 0x0ce148c3u, 0xfb90a602u, 0x819caa36u,
        0x5301fa84u, 0x563d4475u, 0x0dc83629u };
    // 2^256 mod m in 29-bit limbs: mul(x * 2^261, FROM261) = x * 2^256
    static constexpr uint32_t FROM261[9] = { 0x058f0d9du, 0x1aea1c6eu, 0x11c2cf74u, 0x11d651ebu, 0x1462c0a7u,
        0x11b7bc3cu, 0x1cbd99bau, 0x183340fbu, 0x000e0a77u };
    static constexpr uint32_t C783[9] = { 0x0e2312b2u, 0x16c05ca2u, 0x0bc84389u, 0x1cdf310bu, 0x11adafddu,
        0x032e568eu, 0x1d6ae48cu, 0x10d4cd1fu, 0x0026c2d2u };
};

template <class P>
struct F29 {
    static constexpr uint32_t MASK = 0x1fffffffu;
    uint32_t l[9];

    __device__ __forceinline__ static F29 zero()
    {
        F29 r;
#pragma unroll
        for (int i = 0; i < 9; ++i)
            r.l[i] = 0;
        return r;
    }
    __device__ __forceinline__ static F29 one()
    {
        F29 r;
#pragma unroll
        for (int i = 0; i < 9; ++i)
            r.l[i] = P::ONE[i];
        return r;
    }
    __device__ __forceinline__ static F29 from_limbs(const uint32_t (&c)[9])
    {
        F29 r;
#pragma unroll
        for (int i = 0; i < 9; ++i)
            r.l[i] = c[i];
        return r;
    }

    // 8 x 32-bit little-endian words -> 9 x 29-bit limbs (same integer)
    __device__ __forceinline__ static F29 from32(const uint32_t (&v)[8])
    {
        F29 r;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int w = (29 * i) >> 5, s = (29 * i) & 31;
            r.l[i] = __builtin_amdgcn_alignbit(v[w + 1], v[w], s) & MASK;
        }
        r.l[8] = v[7] >> 8;
        return r;
    }
    __device__ __forceinline__ static F29 load32(const void* p)
    {
        const uint4* q = (const uint4*)p;
        const uint4 a = q[0], b = q[1];
        const uint32_t v[8] = { a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w };
        return from32(v);
    }
    __device__ __forceinline__ static F29 from_fp(const Fp256<typename P::Base>& x) { return from32(x.v); }

    // normalised limbs of a value < 2^256 -> 8 x 32-bit words
    __device__ __forceinline__ void to32(uint32_t (&v)[8]) const
    {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int i = (32 * j) / 29, s = 32 * j - 29 * i; // s <= 21: two limbs cover the word
            v[j] = (l[i] >> s) | (l[i + 1] << (29 - s));
        }
    }
    __device__ __forceinline__ void store32(void* p) const
    {
        uint32_t v[8];
        to32(v);
        uint4* q = (uint4*)p;
        q[0] = make_uint4(v[0], v[1], v[2], v[3]);
        q[1] = make_uint4(v[4], v[5], v[6], v[7]);
    }

    // 32 a as an integer, normalised limbs: a normalised word x * 2^256 (value < 2^256)
    // becomes x * 2^261, the same element in the 2^261 scale (value < 2^261), so that
    // mul(shl5(a), b) of two stored words is the stored word of their product
    __device__ __forceinline__ static F29 shl5(const F29& a)
    {
        F29 r;
        r.l[0] = (a.l[0] << 5) & MASK;
#pragma unroll
        for (int i = 1; i < 9; ++i)
            r.l[i] = ((a.l[i] << 5) & MASK) | (a.l[i - 1] >> 24);
        return r;
    }

    // carry limbs 0..7 into their successors (value unchanged)
    __device__ __forceinline__ static F29 normalize(const F29& a)
    {
        F29 r;
        uint32_t c = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const uint32_t t = a.l[i] + c;
            r.l[i] = t & MASK;
            c = t >> 29;
        }
        r.l[8] = a.l[8] + c;
        return r;
    }

    __device__ __forceinline__ static F29 add(const F29& a, const F29& b)
    {
        F29 r;
#pragma unroll
        for (int i = 0; i < 9; ++i)
            r.l[i] = a.l[i] + b.l[i];
        return r;
    }

    // a + 2m - b, limb by limb; b normalised with b < 2m
    __device__ __forceinline__ static F29 sub2(const F29& a, const F29& b)
    {
        F29 r;
#pragma unroll
        for (int i = 0; i < 9; ++i)
            r.l[i] = a.l[i] + P::K2M[i] - b.l[i];
        return r;
    }

    __device__ __forceinline__ static F29 mul(const F29& a, const F29& b)
    {
        uint32_t m[9];
        uint64_t acc = 0;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
#pragma unroll
            for (int i = 0; i <= k; ++i)
                acc += (uint64_t)a.l[i] * b.l[k - i];
#pragma unroll
            for (int i = 0; i < k; ++i)
                acc += (uint64_t)m[i] * P::M[k - i];
            m[k] = ((uint32_t)acc * P::PINV) & MASK;
            acc += (uint64_t)m[k] * P::M[0];
            acc >>= 29;
        }
        F29 r;
#pragma unroll
        for (int k = 9; k < 17; ++k) {
#pragma unroll
            for (int i = k - 8; i <= 8; ++i) {
                acc += (uint64_t)a.l[i] * b.l[k - i];
                acc += (uint64_t)m[i] * P::M[k - i];
            }
            r.l[k - 9] = (uint32_t)acc & MASK;
            acc >>= 29;
        }
        r.l[8] = (uint32_t)acc;
        return r;
    }

    // (a b + c d) / 2^261 with one reduction for both products: every operand
    // normalised (limbs < 2^29, values < 2^259), so a column holds at most 27 products
    // of < 2^58 (< 2^63). Returns normalised limbs of a value < m + (a b + c d) / 2^261.
    __device__ __forceinline__ static F29 mul_sum(const F29& a, const F29& b, const F29& c, const F29& d)
    {
        uint32_t m[9];
        uint64_t acc = 0;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
#pragma unroll
            for (int i = 0; i <= k; ++i) {
                acc += (uint64_t)a.l[i] * b.l[k - i];
                acc += (uint64_t)c.l[i] * d.l[k - i];
            }
#pragma unroll
            for (int i = 0; i < k; ++i)
                acc += (uint64_t)m[i] * P::M[k - i];
            m[k] = ((uint32_t)acc * P::PINV) & MASK;
            acc += (uint64_t)m[k] * P::M[0];
            acc >>= 29;
        }
        F29 r;
#pragma unroll
        for (int k = 9; k < 17; ++k) {
#pragma unroll
            for (int i = k - 8; i <= 8; ++i) {
                acc += (uint64_t)a.l[i] * b.l[k - i];
                acc += (uint64_t)c.l[i] * d.l[k - i];
                acc += (uint64_t)m[i] * P::M[k - i];
            }
            r.l[k - 9] = (uint32_t)acc & MASK;
            acc >>= 29;
        }
        r.l[8] = (uint32_t)acc;
        return r;
    }

    // mul(a, a) with the 36 cross products taken once against doubled limbs
    // (45 instead of 81 multiply-adds for the product half). Same operand
    // limits as mul with both operands a (limbs < 2^30), same result: every
    // column holds the same integer sum, accumulated in the same order of
    // reduction steps.
    __device__ __forceinline__ static F29 sqr(const F29& a)
    {
        uint32_t d[9];
#pragma unroll
        for (int i = 0; i < 9; ++i)
            d[i] = a.l[i] << 1;
        uint32_t m[9];
        uint64_t acc = 0;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
#pragma unroll
            for (int i = 0; 2 * i < k; ++i)
                acc += (uint64_t)d[i] * a.l[k - i];
            if ((k & 1) == 0)
                acc += (uint64_t)a.l[k / 2] * a.l[k / 2];
#pragma unroll
            for (int i = 0; i < k; ++i)
                acc += (uint64_t)m[i] * P::M[k - i];
            m[k] = ((uint32_t)acc * P::PINV) & MASK;
            acc += (uint64_t)m[k] * P::M[0];
            acc >>= 29;
        }
        F29 r;
#pragma unroll
        for (int k = 9; k < 17; ++k) {
#pragma unroll
            for (int i = k - 8; 2 * i < k; ++i)
                acc += (uint64_t)d[i] * a.l[k - i];
            if ((k & 1) == 0)
                acc += (uint64_t)a.l[k / 2] * a.l[k / 2];
#pragma unroll
            for (int i = k - 8; i <= 8; ++i)
                acc += (uint64_t)m[i] * P::M[k - i];
            r.l[k - 9] = (uint32_t)acc & MASK;
            acc >>= 29;
        }
        r.l[8] = (uint32_t)acc;
        return r;
    }

    // a + 4m - b, limb by limb; b normalised with b < 4m
    __device__ __forceinline__ static F29 sub4(const F29& a, const F29& b)
    {
        F29 r;
#pragma unroll
        for (int i = 0; i < 9; ++i)
            r.l[i] = a.l[i] + P::K4M[i] - b.l[i];
        return r;
    }

    // normalised input of value < 2^261 -> normalised representative < 2m
    __device__ __forceinline__ static F29 reduce_2m(const F29& a)
    {
        // quotient estimate from the top 58 bits: q_est in {q - 1, q}
        const double h = (double)a.l[8] * 536870912.0 + (double)a.l[7]; // value / 2^203, truncated
        // 2^203 / m rounded low (m / 2^203 < M[8] * 2^29 + M[7] + 1)
        const double inv = 1.0 / ((double)P::M[8] * 536870912.0 + (double)P::M[7] + 1.0);
        double f = h * inv - 0x1p-30;
        f = f > 0.0 ? f : 0.0;
        const uint32_t q = (uint32_t)f;
        F29 r;
        int64_t c = 0;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            c += (int64_t)a.l[i] - (int64_t)((uint64_t)q * P::M[i]);
            r.l[i] = (uint32_t)c & MASK;
            c >>= 29;
        }
        return r;
    }

    // normalised r < 2m: is r = 0 mod m (r == 0 or r == m)?
    __device__ __forceinline__ static bool is_zero_mod(const F29& r)
    {
        uint32_t z = 0, e = 0;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            z |= r.l[i];
            e |= r.l[i] ^ P::M[i];
        }
        return z == 0 || e == 0;
    }

    __device__ __forceinline__ bool is_zero() const
    {
        uint32_t z = 0;
#pragma unroll
        for (int i = 0; i < 9; ++i)
            z |= l[i];
        return z == 0;
    }

    // a^(m - 2) with Montgomery products (2-bit fixed windows, a small register
    // footprint: it runs in one thread of kernels that keep many values live): for
    // a = u * 2^261 this is u^-1 * 2^261; zero maps to zero. Input normalised, < 2m.
    // Inlined, with the operand by value (callers pass LDS-resident values).
    __device__ __forceinline__ static F29 pow_m_minus_2(const F29 a)
    {
        const F29 a2 = sqr(a);
        const F29 a3 = mul(a2, a);
        uint32_t e[8];
#pragma unroll
        for (int i = 0; i < 8; ++i)
            e[i] = P::Base::M[i];
        e[0] -= 2u; // m is odd and > 2, no borrow
        F29 r = one();
        bool started = false;
        for (int w = 127; w >= 0; --w) {
            const uint32_t d = (e[w >> 4] >> ((w & 15) * 2)) & 3u;
            if (started) {
                r = sqr(r);
                r = sqr(r);
            }
            if (d) {
                const F29& s = d == 1 ? a : (d == 2 ? a2 : a3);
                r = started ? mul(r, s) : s;
                started = true;
            }
        }
        return r;
    }

    // The same value as pow_m_minus_2 (u^-1 * 2^261 for a = u * 2^261, i.e.
    // 2^522 / A for the integer A = a mod m; zero maps to zero) by the binary
    // extended GCD (inv_mod_plain) and one product with 2^783. Input normalised,
    // < 2m; output a normalised representative < 1.01m.
    __device__ static F29 inverse_vt(const F29 a)
    {
        const F29 c = reduce_full(a);
        if (c.is_zero())
            return zero();
        uint32_t w[8], rw[8];
        c.to32(w);
        inv_mod_plain<typename P::Base>(w, rw);
        return mul(from32(rw), from_limbs(P::C783));
    }

    // normalised input of value < 2^261 -> canonical representative in [0, m)
    __device__ __forceinline__ static F29 reduce_full(const F29& a)
    {
        const F29 r = reduce_2m(a);
        // r < 2m: one conditional subtraction
        F29 t;
        int32_t b = 0;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const int32_t d = (int32_t)r.l[i] - (int32_t)P::M[i] + b;
            t.l[i] = (uint32_t)d & MASK;
            b = d >> 29;
        }
        if (b >= 0)
            return t;
        return r;
    }
};

struct Fr29P : Fr29Params {
    using Base = FrParams;
};
struct Fq29P : Fq29Params {
    using Base = FqParams;
};
using Fr29 = F29<Fr29P>;
using Fq29 = F29<Fq29P>;

} // namespace h2
