// BN254 prime-field arithmetic for gfx950 (CDNA4).
//
// Storage format (the boundary format of the reference's FFI, halo2curves'
// in-memory repr): 32 bytes, little-endian, Montgomery form a * 2^256 mod m,
// fully reduced. In registers an element is 8 x 32-bit limbs.
//
// Multiplication is Montgomery product scanning with interleaved reduction
// (FIPS, Koc-Acar-Kaliski 1996): every 32x32 partial product goes through one
// `v_mad_u64_u32` into a 64-bit column accumulator whose carry-out feeds a
// third accumulator word through `v_addc_co_u32`. This keeps the integer
// multiply-add unit (64 lanes/clk/CU on gfx950) the only hot resource; no MFMA
// because the arithmetic is exact modular integer work.
//
// Constants (moduli, -m^-1 mod 2^32, R mod m, R^2 mod m) are BN254's; the
// reference keeps the same values at cuda/include/field/alt_bn128.hpp:19-47.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace h2 {

struct FrParams {
    static constexpr uint32_t M[8] = { 0xf0000001u, 0x43e1f593u, 0x79b97091u, 0x2833e848u,
        0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u };
    static constexpr uint32_t INV = 0xefffffffu;
    static constexpr uint32_t ONE[8] = { 0x4ffffffbu, 0xac96341cu, 0x9f60cd29u, 0x36fc7695u,
        0x7879462eu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u };
    static constexpr uint32_t R2[8] = { 0xae216da7u, 0x1bb8e645u, 0xe35c59e3u, 0x53fe3ab1u,
        0x53bb8085u, 0x8c49833du, 0x7f4e44a5u, 0x0216d0b1u };
    static constexpr uint32_t R3[8] = { 0xb4bf0040u, 0x5e94d8e1u, 0x1cfbb6b8u, 0x2a489cbeu, 0xa19fcfedu,
        0x893cc664u, 0x7fcc657cu, 0x0cf8594bu }; // 2^768 mod m
};

struct FqParams {
    static constexpr uint32_t M[8] = { 0xd87cfd47u, 0x3c208c16u, 0x6871ca8du, 0x97816a91u,
        0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u };
    static constexpr uint32_t INV = 0xe4866389u;
    static constexpr uint32_t ONE[8] = { 0xc58f0d9du, 0xd35d438du, 0xf5c70b3du, 0x0a78eb28u,
        0x7879462cu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u };
    static constexpr uint32_t R2[8] = { 0x538afa89u, 0xf32cfc5bu, 0xd44501fbu, 0xb5e71911u,
        0x0a417ff6u, 0x47ab1effu, 0xcab8351fu, 0x06d89f71u };
    static constexpr uint32_t R3[8] = { 0xda1530dfu, 0xb1cd6dafu, 0xa7283db6u, 0x62f210e6u, 0x0ada0afbu,
        0xef7f0b0cu, 0x2d592544u, 0x20fd6e90u }; // 2^768 mod m
};

// acc(64) + x*y -> acc, carry-out of the 64-bit sum added into c2.
__device__ __forceinline__ void mac96(uint64_t& acc, uint32_t& c2, uint32_t x, uint32_t y)
{
    uint64_t cm;
    asm("v_mad_u64_u32 %0, %1, %3, %4, %0\n\t"
        "v_addc_co_u32_e64 %2, %1, %2, 0, %1"
        : "+v"(acc), "=&s"(cm), "+v"(c2)
        : "v"(x), "v"(y));
}

// x^-1 mod m for 0 < x < m < 2^254, m odd (4 little-endian 64-bit limbs), by the
// binary extended Euclidean algorithm: u = x, v = m, x1 = 1, x2 = 0 with
// x1 x = u and x2 x = v (mod m); halve the even one of u, v (and its x mod m),
// subtract the smaller from the larger, until u or v is 1. Variable time (the
// data is public) and a chain of a few thousand 64-bit operations, against the
// ~380 dependent Montgomery products of an exponentiation.
__device__ inline void inv_mod_binary(const uint64_t (&x)[4], const uint64_t (&m)[4], uint64_t (&out)[4])
{
    uint64_t u[4] = { x[0], x[1], x[2], x[3] }, v[4] = { m[0], m[1], m[2], m[3] };
    uint64_t x1[4] = { 1, 0, 0, 0 }, x2[4] = { 0, 0, 0, 0 };
    auto is_one = [](const uint64_t (&a)[4]) { return a[0] == 1 && (a[1] | a[2] | a[3]) == 0; };
    auto add = [](uint64_t (&a)[4], const uint64_t (&b)[4]) { // no carry out: a + b < 2^256
        uint64_t c = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint64_t s = a[i] + b[i];
            const uint64_t t = s + c;
            c = (uint64_t)(s < a[i]) + (uint64_t)(t < s);
            a[i] = t;
        }
    };
    auto sub = [](uint64_t (&a)[4], const uint64_t (&b)[4]) -> bool { // a -= b, returns the borrow
        uint64_t c = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint64_t d = a[i] - b[i];
            const uint64_t t = d - c;
            c = (uint64_t)(a[i] < b[i]) + (uint64_t)(d < c);
            a[i] = t;
        }
        return c != 0;
    };
    auto geq = [](const uint64_t (&a)[4], const uint64_t (&b)[4]) {
        for (int i = 3; i >= 0; --i)
            if (a[i] != b[i])
                return a[i] > b[i];
        return true;
    };
    auto shr = [](uint64_t (&a)[4], int k) { // 0 < k < 64
        a[0] = (a[0] >> k) | (a[1] << (64 - k));
        a[1] = (a[1] >> k) | (a[2] << (64 - k));
        a[2] = (a[2] >> k) | (a[3] << (64 - k));
        a[3] >>= k;
    };
    auto halve = [&](uint64_t (&a)[4], uint64_t (&t)[4]) { // a >>= k (a even), t = t / 2^k mod m
        const int k = a[0] ? __builtin_ctzll(a[0]) : 63;
        shr(a, k);
        for (int j = 0; j < k; ++j) {
            if (t[0] & 1)
                add(t, m); // t + m < 2m < 2^255
            shr(t, 1);
        }
    };
    while (!is_one(u) && !is_one(v)) {
        while (!(u[0] & 1))
            halve(u, x1);
        while (!(v[0] & 1))
            halve(v, x2);
        if (geq(u, v)) {
            sub(u, v);
            if (sub(x1, x2))
                add(x1, m);
        } else {
            sub(v, u);
            if (sub(x2, x1))
                add(x2, m);
        }
    }
    const uint64_t (&r)[4] = is_one(u) ? x1 : x2;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        out[i] = r[i];
}

template <class P>
struct Fp256 {
    uint32_t v[8];

    __host__ __device__ __forceinline__ static Fp256 zero()
    {
        Fp256 r;
#pragma unroll
        for (int i = 0; i < 8; ++i)
            r.v[i] = 0;
        return r;
    }
    __host__ __device__ __forceinline__ static Fp256 one()
    {
        Fp256 r;
#pragma unroll
        for (int i = 0; i < 8; ++i)
            r.v[i] = P::ONE[i];
        return r;
    }
    __host__ __device__ __forceinline__ static Fp256 r2()
    {
        Fp256 r;
#pragma unroll
        for (int i = 0; i < 8; ++i)
            r.v[i] = P::R2[i];
        return r;
    }
    __device__ __forceinline__ static Fp256 load(const void* p)
    {
        const uint4* q = (const uint4*)p;
        uint4 a = q[0], b = q[1];
        Fp256 r;
        r.v[0] = a.x; r.v[1] = a.y; r.v[2] = a.z; r.v[3] = a.w;
        r.v[4] = b.x; r.v[5] = b.y; r.v[6] = b.z; r.v[7] = b.w;
        return r;
    }
    __device__ __forceinline__ void store(void* p) const
    {
        uint4* q = (uint4*)p;
        q[0] = make_uint4(v[0], v[1], v[2], v[3]);
        q[1] = make_uint4(v[4], v[5], v[6], v[7]);
    }
    __device__ __forceinline__ bool is_zero() const
    {
        uint32_t o = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i)
            o |= v[i];
        return o == 0;
    }
    __device__ __forceinline__ bool operator==(const Fp256& b) const
    {
        uint32_t o = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i)
            o |= v[i] ^ b.v[i];
        return o == 0;
    }
    __device__ __forceinline__ bool operator!=(const Fp256& b) const { return !(*this == b); }

    // r = t - M if t >= M else t   (t < 2M)
    __device__ __forceinline__ static Fp256 reduce_once(const uint32_t* t)
    {
        uint32_t d[8];
        uint32_t borrow = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            uint64_t x = (uint64_t)t[j] - P::M[j] - borrow;
            d[j] = (uint32_t)x;
            borrow = (uint32_t)(x >> 32) & 1u;
        }
        Fp256 r;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            r.v[j] = borrow ? t[j] : d[j];
        return r;
    }

    __device__ __forceinline__ static Fp256 add(const Fp256& a, const Fp256& b)
    {
        uint32_t t[8];
        uint64_t c = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            c = (uint64_t)a.v[j] + b.v[j] + (c >> 32);
            t[j] = (uint32_t)c;
        }
        // a, b < M < 2^254: the sum never leaves 256 bits
        return reduce_once(t);
    }

    __device__ __forceinline__ static Fp256 sub(const Fp256& a, const Fp256& b)
    {
        uint32_t t[8];
        uint32_t borrow = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            uint64_t x = (uint64_t)a.v[j] - b.v[j] - borrow;
            t[j] = (uint32_t)x;
            borrow = (uint32_t)(x >> 32) & 1u;
        }
        uint32_t mask = 0u - borrow;
        Fp256 r;
        uint64_t c = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            c = (uint64_t)t[j] + (P::M[j] & mask) + (c >> 32);
            r.v[j] = (uint32_t)c;
        }
        return r;
    }

    __device__ __forceinline__ static Fp256 neg(const Fp256& a)
    {
        return sub(zero(), a);
    }

    __device__ __forceinline__ static Fp256 mul(const Fp256& a, const Fp256& b)
    {
        uint32_t m[8];
        uint32_t t[8];
        uint64_t acc = 0;
        uint32_t c2 = 0;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int j = k - i;
                if (j >= 0 && j < 8)
                    mac96(acc, c2, a.v[i], b.v[j]);
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int j = k - i;
                if (i < k && j >= 1 && j < 8)
                    mac96(acc, c2, m[i], P::M[j]);
            }
            if (k < 8) {
                m[k] = (uint32_t)acc * P::INV;
                mac96(acc, c2, m[k], P::M[0]);
            } else {
                t[k - 8] = (uint32_t)acc;
            }
            acc = (acc >> 32) | ((uint64_t)c2 << 32);
            c2 = 0;
        }
        // M < 2^254 < R/4 keeps the FIPS output below 2M
        return reduce_once(t);
    }

    __device__ __forceinline__ static Fp256 sqr(const Fp256& a) { return mul(a, a); }

    __device__ __forceinline__ static Fp256 dbl(const Fp256& a) { return add(a, a); }

    // a^e for a small exponent (square-and-multiply, msb first)
    __device__ __forceinline__ static Fp256 pow_u64(const Fp256& a, uint64_t e)
    {
        Fp256 r = one();
        if (e == 0)
            return r;
        int top = 63 - __clzll(e);
        for (int i = top; i >= 0; --i) {
            r = sqr(r);
            if ((e >> i) & 1ull)
                r = mul(r, a);
        }
        return r;
    }

    // a^-1 by the binary extended Euclidean algorithm: plain inverse of the
    // stored word A = a R, times R^3 with one Montgomery product (A^-1 R^2 = a^-1 R);
    // inverse_vt(0) = 0
    __device__ static Fp256 inverse_vt(const Fp256& a)
    {
        if (a.is_zero())
            return zero();
        uint64_t x[4], m[4], r[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            x[i] = ((uint64_t)a.v[2 * i + 1] << 32) | a.v[2 * i];
            m[i] = ((uint64_t)P::M[2 * i + 1] << 32) | P::M[2 * i];
        }
        inv_mod_binary(x, m, r);
        Fp256 ri, r3;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            ri.v[2 * i] = (uint32_t)r[i];
            ri.v[2 * i + 1] = (uint32_t)(r[i] >> 32);
        }
#pragma unroll
        for (int i = 0; i < 8; ++i)
            r3.v[i] = P::R3[i];
        return mul(ri, r3);
    }

    // a^(M-2) with 4-bit fixed windows; inverse(0) = 0
    __device__ static Fp256 inverse(const Fp256& a)
    {
        Fp256 tbl[16];
        tbl[0] = one();
        tbl[1] = a;
#pragma unroll
        for (int i = 2; i < 16; ++i)
            tbl[i] = mul(tbl[i - 1], a);
        uint32_t e[8];
#pragma unroll
        for (int i = 0; i < 8; ++i)
            e[i] = P::M[i];
        e[0] -= 2u; // M is odd and > 2, no borrow
        Fp256 r = one();
        bool started = false;
        for (int w = 63; w >= 0; --w) {
            uint32_t nib = (e[w >> 3] >> ((w & 7) * 4)) & 0xfu;
            if (started) {
                r = sqr(r);
                r = sqr(r);
                r = sqr(r);
                r = sqr(r);
            }
            if (nib) {
                // select without dynamic register indexing
                Fp256 s = tbl[0];
#pragma unroll
                for (int i = 1; i < 16; ++i)
                    if ((uint32_t)i == nib)
                        s = tbl[i];
                r = started ? mul(r, s) : s;
                started = true;
            }
        }
        return r;
    }

    __device__ __forceinline__ static Fp256 from_mont(const Fp256& a)
    {
        Fp256 o = zero();
        o.v[0] = 1;
        return mul(a, o);
    }
    __device__ __forceinline__ static Fp256 to_mont(const Fp256& a) { return mul(a, r2()); }
};

template <class P>
__device__ __forceinline__ Fp256<P> operator*(const Fp256<P>& a, const Fp256<P>& b) { return Fp256<P>::mul(a, b); }
template <class P>
__device__ __forceinline__ Fp256<P> operator+(const Fp256<P>& a, const Fp256<P>& b) { return Fp256<P>::add(a, b); }
template <class P>
__device__ __forceinline__ Fp256<P> operator-(const Fp256<P>& a, const Fp256<P>& b) { return Fp256<P>::sub(a, b); }

using Fr = Fp256<FrParams>;
using Fq = Fp256<FqParams>;

} // namespace h2
