// BN254 G1 (y^2 = x^3 + 3 over Fq) on gfx950.
//
// Buckets accumulate in extended Jacobian "XYZZ" coordinates (x = X/ZZ,
// y = Y/ZZZ, ZZ^3 = ZZZ^2); the identity is ZZ = ZZZ = 0. Formulas are the
// EFD's for short Weierstrass a = 0:
//   madd-2008-s (XYZZ + affine, 8M + 2S), add-2008-s (12M + 2S),
//   mdbl-2008-s / dbl-2008-s for the doubling cases.
// Affine points are 64 bytes (x, y), the identity is (0, 0).
#pragma once

#include "field.h"

namespace h2 {

struct G1Affine {
    Fq x, y;

    __device__ __forceinline__ static G1Affine load(const void* p)
    {
        const uint8_t* b = (const uint8_t*)p;
        return G1Affine { Fq::load(b), Fq::load(b + 32) };
    }
    __device__ __forceinline__ void store(void* p) const
    {
        uint8_t* b = (uint8_t*)p;
        x.store(b);
        y.store(b + 32);
    }
    __device__ __forceinline__ bool is_inf() const { return x.is_zero() && y.is_zero(); }
    __device__ __forceinline__ G1Affine neg() const { return is_inf() ? *this : G1Affine { x, Fq::neg(y) }; }
};

struct G1Xyzz {
    Fq X, Y, ZZ, ZZZ;

    __device__ __forceinline__ static G1Xyzz inf()
    {
        G1Xyzz r;
        r.X = Fq::one();
        r.Y = Fq::one();
        r.ZZ = Fq::zero();
        r.ZZZ = Fq::zero();
        return r;
    }
    __device__ __forceinline__ bool is_inf() const { return ZZ.is_zero(); }

    __device__ __forceinline__ static G1Xyzz from_affine(const G1Affine& a)
    {
        if (a.is_inf())
            return inf();
        return G1Xyzz { a.x, a.y, Fq::one(), Fq::one() };
    }

    __device__ __forceinline__ static G1Xyzz load(const void* p)
    {
        const uint8_t* b = (const uint8_t*)p;
        return G1Xyzz { Fq::load(b), Fq::load(b + 32), Fq::load(b + 64), Fq::load(b + 96) };
    }
    __device__ __forceinline__ void store(void* p) const
    {
        uint8_t* b = (uint8_t*)p;
        X.store(b);
        Y.store(b + 32);
        ZZ.store(b + 64);
        ZZZ.store(b + 96);
    }

    // mdbl-2008-s: 2 * (x, y) for an affine, non-identity point
    __device__ __noinline__ static G1Xyzz dbl_affine(const G1Affine& a)
    {
        Fq U = Fq::dbl(a.y);
        Fq V = Fq::sqr(U);
        Fq W = U * V;
        Fq S = a.x * V;
        Fq x2 = Fq::sqr(a.x);
        Fq M = Fq::dbl(x2) + x2;
        G1Xyzz r;
        r.X = Fq::sqr(M) - Fq::dbl(S);
        r.Y = M * (S - r.X) - W * a.y;
        r.ZZ = V;
        r.ZZZ = W;
        return r;
    }

    // dbl-2008-s
    __device__ __noinline__ static G1Xyzz dbl(const G1Xyzz& p)
    {
        if (p.is_inf())
            return p;
        Fq U = Fq::dbl(p.Y);
        Fq V = Fq::sqr(U);
        Fq W = U * V;
        Fq S = p.X * V;
        Fq x2 = Fq::sqr(p.X);
        Fq M = Fq::dbl(x2) + x2;
        G1Xyzz r;
        r.X = Fq::sqr(M) - Fq::dbl(S);
        r.Y = M * (S - r.X) - W * p.Y;
        r.ZZ = V * p.ZZ;
        r.ZZZ = W * p.ZZZ;
        return r;
    }

    // madd-2008-s with the exceptional cases: p + q, q affine
    __device__ static G1Xyzz add_affine(const G1Xyzz& p, const G1Affine& q)
    {
        if (q.is_inf())
            return p;
        if (p.is_inf())
            return from_affine(q);
        Fq U2 = q.x * p.ZZ;
        Fq S2 = q.y * p.ZZZ;
        Fq P = U2 - p.X;
        Fq R = S2 - p.Y;
        if (P.is_zero()) {
            if (R.is_zero())
                return dbl_affine(q);
            return inf();
        }
        Fq PP = Fq::sqr(P);
        Fq PPP = P * PP;
        Fq Q = p.X * PP;
        G1Xyzz r;
        r.X = Fq::sqr(R) - PPP - Fq::dbl(Q);
        r.Y = R * (Q - r.X) - p.Y * PPP;
        r.ZZ = p.ZZ * PP;
        r.ZZZ = p.ZZZ * PPP;
        return r;
    }

    // add-2008-s with the exceptional cases
    __device__ static G1Xyzz add(const G1Xyzz& p, const G1Xyzz& q)
    {
        if (q.is_inf())
            return p;
        if (p.is_inf())
            return q;
        Fq U1 = p.X * q.ZZ;
        Fq U2 = q.X * p.ZZ;
        Fq S1 = p.Y * q.ZZZ;
        Fq S2 = q.Y * p.ZZZ;
        Fq P = U2 - U1;
        Fq R = S2 - S1;
        if (P.is_zero()) {
            if (R.is_zero())
                return dbl(p);
            return inf();
        }
        Fq PP = Fq::sqr(P);
        Fq PPP = P * PP;
        Fq Q = U1 * PP;
        G1Xyzz r;
        r.X = Fq::sqr(R) - PPP - Fq::dbl(Q);
        r.Y = R * (Q - r.X) - S1 * PPP;
        r.ZZ = p.ZZ * q.ZZ * PP;
        r.ZZZ = p.ZZZ * q.ZZZ * PPP;
        return r;
    }

    // k * p for a small k (double-and-add, msb first)
    __device__ static G1Xyzz mul_small(const G1Xyzz& p, uint32_t k)
    {
        G1Xyzz acc = inf();
        if (k == 0 || p.is_inf())
            return acc;
        int top = 31 - __clz(k);
        acc = p;
        for (int i = top - 1; i >= 0; --i) {
            acc = dbl(acc);
            if ((k >> i) & 1u)
                acc = add(acc, p);
        }
        return acc;
    }

    // affine (x, y); identity -> (0, 0)
    __device__ G1Affine to_affine() const
    {
        if (is_inf())
            return G1Affine { Fq::zero(), Fq::zero() };
        const Fq t = Fq::inverse_vt(ZZ * ZZZ); // 1 / (ZZ ZZZ)
        return G1Affine { X * (t * ZZZ), Y * (t * ZZ) };
    }
};

} // namespace h2
