// BN254 G1 bucket arithmetic on the 29-bit-limb field form (field29.h).
//
// Same formulas as curve.h (EFD, short Weierstrass a = 0, XYZZ coordinates:
// madd-2008-s, add-2008-s, dbl-2008-s, mdbl-2008-s) with lazy reduction.
// Every coordinate lives in the 2^261 Montgomery scale; the MSM converts the
// bases into it once per call and converts the result back.
//
// State invariant: the four coordinates of a G1X29 are normalised with values
// < 2m (X is brought below 2m by reduce_2m at the end of a formula; Y is a sum of
// two products under one reduction, mul_sum, < 1.2m; ZZ and ZZZ are products).
// Affine inputs are canonical (< m). With inputs below 2m every product below is
// < 1.06m, so each difference fits the 2m / 4m offsets of sub2 / sub4; the
// comments give the bounds (in units of m).
#pragma once

#include "curve.h"
#include "field29.h"

namespace h2 {

using Fq29 = F29<Fq29P>;

__device__ __forceinline__ Fq29 nsub2(const Fq29& a, const Fq29& b) { return Fq29::normalize(Fq29::sub2(a, b)); }
// 4m - b, normalised (b normalised, < 4m): the negation mul_sum takes for a difference of products
__device__ __forceinline__ Fq29 nneg4(const Fq29& b) { return Fq29::normalize(Fq29::sub4(Fq29::zero(), b)); }

struct G1A29 {
    Fq29 x, y;

    __device__ __forceinline__ static G1A29 load(const void* p)
    {
        const uint8_t* b = (const uint8_t*)p;
        return G1A29 { Fq29::load32(b), Fq29::load32(b + 32) };
    }
    __device__ __forceinline__ bool is_inf() const { return x.is_zero() && y.is_zero(); }
    // -(x, y) = (x, 2m - y): y stays < 2m, normalised (G1 has no 2-torsion, y != 0)
    __device__ __forceinline__ G1A29 neg() const { return is_inf() ? *this : G1A29 { x, nsub2(Fq29::zero(), y) }; }
};

struct G1X29 {
    Fq29 X, Y, ZZ, ZZZ;

    __device__ __forceinline__ static G1X29 inf()
    {
        return G1X29 { Fq29::one(), Fq29::one(), Fq29::zero(), Fq29::zero() };
    }
    __device__ __forceinline__ bool is_inf() const { return ZZ.is_zero(); }

    __device__ __forceinline__ static G1X29 from_affine(const G1A29& a)
    {
        if (a.is_inf())
            return inf();
        return G1X29 { a.x, a.y, Fq29::one(), Fq29::one() };
    }

    // 128 bytes: four coordinates in the 32-byte word layout (values < 2m < 2^256)
    __device__ __forceinline__ static G1X29 load(const void* p)
    {
        const uint8_t* b = (const uint8_t*)p;
        return G1X29 { Fq29::load32(b), Fq29::load32(b + 32), Fq29::load32(b + 64), Fq29::load32(b + 96) };
    }
    __device__ __forceinline__ void store(void* p) const
    {
        uint8_t* b = (uint8_t*)p;
        X.store32(b);
        Y.store32(b + 32);
        ZZ.store32(b + 64);
        ZZZ.store32(b + 96);
    }

    // mdbl-2008-s: 2 (x, y), affine non-identity input
    __device__ __forceinline__ static G1X29 dbl_affine(const G1A29& a)
    {
        const Fq29 U = Fq29::add(a.y, a.y); // < 4, limbs < 2^30
        const Fq29 V = Fq29::sqr(U); // < 1.1
        const Fq29 W = Fq29::mul(U, V);
        const Fq29 S = Fq29::mul(a.x, V);
        const Fq29 x2 = Fq29::sqr(a.x);
        const Fq29 M = Fq29::normalize(Fq29::add(Fq29::add(x2, x2), x2)); // < 3.3
        const Fq29 S2 = Fq29::normalize(Fq29::add(S, S)); // < 2.2
        G1X29 r;
        r.X = Fq29::reduce_2m(Fq29::normalize(Fq29::sub4(Fq29::sqr(M), S2)));
        // M (S - X3) - W y with one reduction: < m + (3.3 * 3.1 + 1.1 * 4) m^2 / 2^261 < 1.2m
        r.Y = Fq29::mul_sum(M, nsub2(S, r.X), W, nneg4(a.y));
        r.ZZ = V;
        r.ZZZ = W;
        return r;
    }

    // dbl-2008-s (out of line: it is the rare branch of add)
    __device__ __noinline__ static G1X29 dbl(const G1X29& p) { return dbl_inline(p); }

    __device__ __forceinline__ static G1X29 dbl_inline(const G1X29& p)
    {
        if (p.is_inf())
            return p;
        const Fq29 U = Fq29::add(p.Y, p.Y); // < 4, limbs < 2^30
        const Fq29 V = Fq29::sqr(U);
        const Fq29 W = Fq29::mul(U, V);
        const Fq29 S = Fq29::mul(p.X, V);
        const Fq29 x2 = Fq29::sqr(p.X);
        const Fq29 M = Fq29::normalize(Fq29::add(Fq29::add(x2, x2), x2));
        const Fq29 S2 = Fq29::normalize(Fq29::add(S, S));
        G1X29 r;
        r.X = Fq29::reduce_2m(Fq29::normalize(Fq29::sub4(Fq29::sqr(M), S2)));
        r.Y = Fq29::mul_sum(M, nsub2(S, r.X), W, nneg4(p.Y)); // < 1.2m, as in dbl_affine
        r.ZZ = Fq29::mul(V, p.ZZ);
        r.ZZZ = Fq29::mul(W, p.ZZZ);
        return r;
    }

    // madd-2008-s with the exceptional cases: p + q, q affine
    __device__ static G1X29 add_affine(const G1X29& p, const G1A29& q)
    {
        if (q.is_inf())
            return p;
        if (p.is_inf())
            return from_affine(q);
        const Fq29 U2 = Fq29::mul(q.x, p.ZZ);
        const Fq29 S2 = Fq29::mul(q.y, p.ZZZ);
        const Fq29 P = Fq29::reduce_2m(nsub2(U2, p.X)); // < 2
        const Fq29 R = nsub2(S2, p.Y); // < 3.1
        if (Fq29::is_zero_mod(P)) {
            if (Fq29::is_zero_mod(Fq29::reduce_2m(R)))
                return dbl_affine(q);
            return inf();
        }
        const Fq29 PP = Fq29::sqr(P);
        const Fq29 PPP = Fq29::mul(P, PP);
        const Fq29 Q = Fq29::mul(p.X, PP);
        const Fq29 T = Fq29::normalize(Fq29::add(PPP, Fq29::add(Q, Q))); // < 3.2
        G1X29 r;
        r.X = Fq29::reduce_2m(Fq29::normalize(Fq29::sub4(Fq29::sqr(R), T)));
        // R (Q - X3) - Y1 PPP with one reduction: < m + (3.1 * 3.1 + 2 * 4) m^2 / 2^261 < 1.2m
        r.Y = Fq29::mul_sum(R, nsub2(Q, r.X), p.Y, nneg4(PPP));
        r.ZZ = Fq29::mul(p.ZZ, PP);
        r.ZZZ = Fq29::mul(p.ZZZ, PPP);
        return r;
    }

    // add-2008-s with the exceptional cases
    __device__ static G1X29 add(const G1X29& p, const G1X29& q)
    {
        if (q.is_inf())
            return p;
        if (p.is_inf())
            return q;
        const Fq29 U1 = Fq29::mul(p.X, q.ZZ);
        const Fq29 U2 = Fq29::mul(q.X, p.ZZ);
        const Fq29 S1 = Fq29::mul(p.Y, q.ZZZ);
        const Fq29 S2 = Fq29::mul(q.Y, p.ZZZ);
        const Fq29 P = Fq29::reduce_2m(nsub2(U2, U1));
        const Fq29 R = nsub2(S2, S1);
        if (Fq29::is_zero_mod(P)) {
            if (Fq29::is_zero_mod(Fq29::reduce_2m(R)))
                return dbl(p);
            return inf();
        }
        const Fq29 PP = Fq29::sqr(P);
        const Fq29 PPP = Fq29::mul(P, PP);
        const Fq29 Q = Fq29::mul(U1, PP);
        const Fq29 T = Fq29::normalize(Fq29::add(PPP, Fq29::add(Q, Q)));
        G1X29 r;
        r.X = Fq29::reduce_2m(Fq29::normalize(Fq29::sub4(Fq29::sqr(R), T)));
        r.Y = Fq29::mul_sum(R, nsub2(Q, r.X), S1, nneg4(PPP)); // < 1.2m, as in add_affine
        r.ZZ = Fq29::mul(Fq29::mul(p.ZZ, q.ZZ), PP);
        r.ZZZ = Fq29::mul(Fq29::mul(p.ZZZ, q.ZZZ), PPP);
        return r;
    }

    // XYZZ -> Jacobian (X ZZ^2, Y ZZZ^2, ZZZ): x = X ZZ^2 / ZZZ^2 = X / ZZ and
    // y = Y ZZZ^2 / ZZZ^3 because ZZ^3 = ZZZ^2; the identity keeps Z = 0
    __device__ __forceinline__ void to_jacobian(Fq29& jx, Fq29& jy, Fq29& jz) const
    {
        jx = Fq29::mul(X, Fq29::sqr(ZZ));
        jy = Fq29::mul(Y, Fq29::sqr(ZZZ));
        jz = ZZZ;
    }
    __device__ __forceinline__ static G1X29 from_jacobian(const Fq29& jx, const Fq29& jy, const Fq29& jz)
    {
        const Fq29 zz = Fq29::sqr(jz);
        return G1X29 { jx, jy, zz, Fq29::mul(zz, jz) };
    }

    // back to the 2^256 scale (canonical) for the final affine conversion
    __device__ G1Xyzz to_xyzz256() const
    {
        const Fq29 f = Fq29::from_limbs(Fq29P::FROM261);
        G1Xyzz r;
        Fq29::reduce_full(Fq29::mul(X, f)).to32(r.X.v);
        Fq29::reduce_full(Fq29::mul(Y, f)).to32(r.Y.v);
        Fq29::reduce_full(Fq29::mul(ZZ, f)).to32(r.ZZ.v);
        Fq29::reduce_full(Fq29::mul(ZZZ, f)).to32(r.ZZZ.v);
        return r;
    }
};

} // namespace h2
