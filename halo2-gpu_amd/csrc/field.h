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

// x^-1 mod m (P::M, odd, < 2^254) for 0 < x < m, both as 8 little-endian 32-bit limbs:
// the binary extended GCD in Pornin's batched form ("Optimized Binary GCD for Modular
// Inversion", 2020, Algorithm 2 with k = 32). The state is a, b (a = x, b = m) and u, v
// with a = u x and b = v x (mod m) up to a known power of 2. Each of 17 outer rounds runs
// 31 binary-GCD steps (halve the even one; when a is odd, swap so that a >= b and subtract)
// on 64-bit images of a and b (their low 31 bits, exact, which fix every parity, and the
// top 33 bits of the longer one, for the comparisons), recording the steps as a 2x2 matrix
// (f0 g0; f1 g1) with |f| + |g| <= 2^31. The matrix is then applied once to the full
// values: a, b <- (f0 a + g0 b) / 2^31, (f1 a + g1 b) / 2^31 (exact; a negative result is
// negated with its row), and u, v <- f0 u + g0 v, f1 u + g1 v, each reduced mod m by one
// Montgomery step of 2^32. 17 x 31 = 527 >= 2 x 254 - 1 steps leave b = gcd = 1 and
// v = x^-1 2^(527 - 544); the final 17 doublings mod m remove the factor. Variable time
// (the data is public): ~17 x 31 cheap 64-bit steps and 68 short multi-precision
// products, against ~700 full-width steps of the plain binary algorithm.
template <class P>
__host__ __device__ inline void inv_mod_plain(const uint32_t (&x)[8], uint32_t (&out)[8])
{
    uint32_t a[8], b[8], u[8], v[8];
    for (int i = 0; i < 8; ++i) {
        a[i] = x[i];
        b[i] = P::M[i];
        u[i] = i == 0 ? 1u : 0u;
        v[i] = 0;
    }
    auto bitlen = [](const uint32_t (&w)[8]) -> int {
        for (int i = 7; i >= 0; --i)
            if (w[i])
                return 32 * i + 32 - __builtin_clz(w[i]);
        return 0;
    };
    // bits [s, s + 33) of w (s >= 31; limbs past the top are zero)
    auto top33 = [](const uint32_t (&w)[8], int s) -> uint64_t {
        const int q = s >> 5, sh = s & 31;
        const uint64_t w0 = q < 8 ? w[q] : 0, w1 = q + 1 < 8 ? w[q + 1] : 0, w2 = q + 2 < 8 ? w[q + 2] : 0;
        uint64_t r = ((w1 << 32) | w0) >> sh;
        if (sh)
            r |= w2 << (64 - sh);
        return r & ((1ull << 33) - 1);
    };
    // r = (f p + g q) as a signed 288-bit value in 9 limbs (two's complement), p, q < 2^256
    auto lincomb = [](const uint32_t (&p)[8], const uint32_t (&q)[8], int64_t f, int64_t g, uint32_t (&r)[9]) {
        const uint64_t fa = (uint64_t)(f < 0 ? -f : f), ga = (uint64_t)(g < 0 ? -g : g); // <= 2^31
        uint64_t cp = 0, cq = 0;
        int64_t c = 0;
        for (int i = 0; i < 9; ++i) {
            uint64_t tp = cp, tq = cq;
            if (i < 8) {
                tp += (uint64_t)p[i] * fa;
                tq += (uint64_t)q[i] * ga;
            }
            const int64_t lp = (int64_t)(uint32_t)tp, lq = (int64_t)(uint32_t)tq;
            cp = tp >> 32;
            cq = tq >> 32;
            const int64_t t = (f < 0 ? -lp : lp) + (g < 0 ? -lq : lq) + c;
            r[i] = (uint32_t)t;
            c = t >> 32; // arithmetic
        }
    };
    // a, b update: r / 2^31 with its sign; returns true when r was negative (then |r| / 2^31)
    auto shift31 = [](uint32_t (&r)[9], uint32_t (&o)[8]) -> bool {
        const bool neg = (int32_t)r[8] < 0;
        if (neg) { // two's complement negation
            uint64_t c = 1;
            for (int i = 0; i < 9; ++i) {
                const uint64_t t = (uint64_t)(uint32_t)~r[i] + c;
                r[i] = (uint32_t)t;
                c = t >> 32;
            }
        }
        for (int i = 0; i < 8; ++i)
            o[i] = (r[i] >> 31) | (r[i + 1] << 1);
        return neg;
    };
    // u, v update: (f p + g q + 2^32 m) / 2^32 mod m (one Montgomery step), p, q < m
    auto modcomb = [&](const uint32_t (&p)[8], const uint32_t (&q)[8], int64_t f, int64_t g, uint32_t (&o)[8]) {
        uint32_t r[9];
        lincomb(p, q, f, g, r); // |r| <= 2^31 m
        uint32_t t[10];
        uint64_t c = 0; // + 2^32 m: nonnegative, below 2^33 m < 2^288
        for (int i = 0; i < 10; ++i) {
            const uint64_t rv = i < 9 ? (uint64_t)r[i] : ((int32_t)r[8] < 0 ? 0xffffffffull : 0ull);
            const uint64_t mv = (i >= 1 && i <= 8) ? (uint64_t)P::M[i - 1] : 0ull;
            const uint64_t s = rv + mv + c;
            t[i] = (uint32_t)s;
            c = s >> 32;
        }
        const uint32_t qm = t[0] * P::INV; // t + qm m = 0 mod 2^32
        c = 0;
        for (int i = 0; i < 10; ++i) {
            const uint64_t s = (uint64_t)t[i] + (i < 8 ? (uint64_t)qm * P::M[i] : 0ull) + c;
            t[i] = (uint32_t)s;
            c = s >> 32;
        }
        for (int i = 0; i < 8; ++i) // (t + qm m) / 2^32 < 3m
            o[i] = t[i + 1];
        for (int k = 0; k < 2; ++k) {
            uint32_t d[8];
            int64_t bw = 0;
            for (int i = 0; i < 8; ++i) {
                const int64_t s = (int64_t)o[i] - (int64_t)P::M[i] + bw;
                d[i] = (uint32_t)s;
                bw = s >> 32;
            }
            if (bw == 0)
                for (int i = 0; i < 8; ++i)
                    o[i] = d[i];
        }
    };
    for (int round = 0; round < 17; ++round) {
        int n = bitlen(a) > bitlen(b) ? bitlen(a) : bitlen(b);
        if (n < 64)
            n = 64;
        uint64_t at = ((uint64_t)a[0] & 0x7fffffffu) | (top33(a, n - 33) << 31);
        uint64_t bt = ((uint64_t)b[0] & 0x7fffffffu) | (top33(b, n - 33) << 31);
        int64_t f0 = 1, g0 = 0, f1 = 0, g1 = 1;
        for (int j = 0; j < 31; ++j) {
            if (at & 1) {
                if (at < bt) {
                    const uint64_t t = at;
                    at = bt;
                    bt = t;
                    int64_t s = f0;
                    f0 = f1;
                    f1 = s;
                    s = g0;
                    g0 = g1;
                    g1 = s;
                }
                at -= bt;
                f0 -= f1;
                g0 -= g1;
            }
            at >>= 1;
            f1 *= 2;
            g1 *= 2;
        }
        uint32_t ra[9], rb[9], na[8], nb[8];
        lincomb(a, b, f0, g0, ra);
        lincomb(a, b, f1, g1, rb);
        if (shift31(ra, na)) {
            f0 = -f0;
            g0 = -g0;
        }
        if (shift31(rb, nb)) {
            f1 = -f1;
            g1 = -g1;
        }
        uint32_t nu[8], nv[8];
        modcomb(u, v, f0, g0, nu);
        modcomb(u, v, f1, g1, nv);
        for (int i = 0; i < 8; ++i) {
            a[i] = na[i];
            b[i] = nb[i];
            u[i] = nu[i];
            v[i] = nv[i];
        }
    }
    // b = 1 = v x 2^(544 - 527): x^-1 = v 2^17 mod m
    for (int k = 0; k < 17; ++k) {
        uint32_t d[8];
        uint32_t c = 0;
        for (int i = 0; i < 8; ++i) { // v < m < 2^254: 2v < 2^255, no carry out
            d[i] = (v[i] << 1) | c;
            c = v[i] >> 31;
        }
        uint32_t e[8];
        int64_t bw = 0;
        for (int i = 0; i < 8; ++i) {
            const int64_t s = (int64_t)d[i] - (int64_t)P::M[i] + bw;
            e[i] = (uint32_t)s;
            bw = s >> 32;
        }
        for (int i = 0; i < 8; ++i)
            v[i] = bw == 0 ? e[i] : d[i];
    }
    for (int i = 0; i < 8; ++i)
        out[i] = v[i];
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

    // a^-1 by the binary extended GCD (inv_mod_plain): plain inverse of the
    // stored word A = a R, times R^3 with one Montgomery product (A^-1 R^2 = a^-1 R);
    // inverse_vt(0) = 0
    __device__ static Fp256 inverse_vt(const Fp256& a)
    {
        if (a.is_zero())
            return zero();
        Fp256 ri, r3;
        inv_mod_plain<P>(a.v, ri.v);
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
