/*
 * CPU oracle for the halo2-axiom-gpu hot path.
 *
 * TEST INFRASTRUCTURE ONLY: tests/, the smoke test and the CPU-baseline leg of
 * bench.py load this library as the checker. The product path never links or
 * calls it.
 *
 * Every routine restates the reference CPU algorithm it cites (paths relative
 * to the reference's halo2_proofs/ directory). Field constants are the BN254
 * moduli and Montgomery constants of cuda/include/field/alt_bn128.hpp:19-47;
 * they are recomputed from first principles by tests/test_oracle_pins.py.
 */
#include "halo2_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

typedef unsigned __int128 u128;

typedef struct {
    uint64_t m[4];
    uint64_t inv; /* -m^{-1} mod 2^64 */
    uint64_t r2[4]; /* 2^512 mod m */
    uint64_t one[4]; /* 2^256 mod m */
} fparams;

static const fparams FR = {
    { 0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL },
    0xc2e1f593efffffffULL,
    { 0x1bb8e645ae216da7ULL, 0x53fe3ab1e35c59e3ULL, 0x8c49833d53bb8085ULL, 0x0216d0b17f4e44a5ULL },
    { 0xac96341c4ffffffbULL, 0x36fc76959f60cd29ULL, 0x666ea36f7879462eULL, 0x0e0a77c19a07df2fULL },
};
static const fparams FQ = {
    { 0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL },
    0x87d20782e4866389ULL,
    { 0xf32cfc5b538afa89ULL, 0xb5e71911d44501fbULL, 0x47ab1eff0a417ff6ULL, 0x06d89f71cab8351fULL },
    { 0xd35d438dc58f0d9dULL, 0x0a78eb28f5c70b3dULL, 0x666ea36f7879462cULL, 0x0e0a77c19a07df2fULL },
};

#define INL static inline __attribute__((always_inline))

/* ------------------------------------------------------------------ field */

INL int f_geq(const uint64_t* a, const uint64_t* m)
{
    for (int i = 3; i >= 0; --i) {
        if (a[i] != m[i])
            return a[i] > m[i];
    }
    return 1;
}

INL int f_is_zero(const uint64_t* a) { return (a[0] | a[1] | a[2] | a[3]) == 0; }
INL int f_eq(const uint64_t* a, const uint64_t* b)
{
    return a[0] == b[0] && a[1] == b[1] && a[2] == b[2] && a[3] == b[3];
}
INL void f_copy(uint64_t* r, const uint64_t* a) { memcpy(r, a, 32); }
INL void f_zero(uint64_t* r) { memset(r, 0, 32); }

INL uint64_t sub4(uint64_t* r, const uint64_t* a, const uint64_t* b)
{
    uint64_t borrow = 0;
    for (int i = 0; i < 4; ++i) {
        u128 d = (u128)a[i] - b[i] - borrow;
        r[i] = (uint64_t)d;
        borrow = (uint64_t)(d >> 64) & 1;
    }
    return borrow;
}

INL void f_add(const fparams* P, uint64_t* r, const uint64_t* a, const uint64_t* b)
{
    uint64_t t[4];
    uint64_t carry = 0;
    for (int i = 0; i < 4; ++i) {
        u128 s = (u128)a[i] + b[i] + carry;
        t[i] = (uint64_t)s;
        carry = (uint64_t)(s >> 64);
    }
    /* a, b < m < 2^254 so the sum never carries out of 256 bits */
    if (f_geq(t, P->m))
        sub4(t, t, P->m);
    f_copy(r, t);
}

INL void f_sub(const fparams* P, uint64_t* r, const uint64_t* a, const uint64_t* b)
{
    uint64_t t[4];
    if (sub4(t, a, b)) {
        uint64_t carry = 0;
        for (int i = 0; i < 4; ++i) {
            u128 s = (u128)t[i] + P->m[i] + carry;
            t[i] = (uint64_t)s;
            carry = (uint64_t)(s >> 64);
        }
    }
    f_copy(r, t);
}

INL void f_dbl(const fparams* P, uint64_t* r, const uint64_t* a) { f_add(P, r, a, a); }

/* Montgomery multiplication, CIOS (Koc, Acar, Kaliski 1996), 4 x 64-bit limbs. */
INL void f_mul(const fparams* P, uint64_t* r, const uint64_t* a, const uint64_t* b)
{
    uint64_t t[6] = { 0, 0, 0, 0, 0, 0 };
    for (int i = 0; i < 4; ++i) {
        u128 c = 0;
        for (int j = 0; j < 4; ++j) {
            c = (u128)a[j] * b[i] + t[j] + (uint64_t)(c >> 64);
            t[j] = (uint64_t)c;
        }
        c = (u128)t[4] + (uint64_t)(c >> 64);
        t[4] = (uint64_t)c;
        t[5] = (uint64_t)(c >> 64);
        uint64_t m = t[0] * P->inv;
        c = (u128)m * P->m[0] + t[0];
        for (int j = 1; j < 4; ++j) {
            c = (u128)m * P->m[j] + t[j] + (uint64_t)(c >> 64);
            t[j - 1] = (uint64_t)c;
        }
        c = (u128)t[4] + (uint64_t)(c >> 64);
        t[3] = (uint64_t)c;
        t[4] = t[5] + (uint64_t)(c >> 64);
    }
    if (t[4] || f_geq(t, P->m))
        sub4(t, t, P->m);
    f_copy(r, t);
}

INL void f_sqr(const fparams* P, uint64_t* r, const uint64_t* a) { f_mul(P, r, a, a); }

static void f_pow_limbs(const fparams* P, uint64_t* r, const uint64_t* a, const uint64_t* e)
{
    uint64_t acc[4], base[4];
    f_copy(acc, P->one);
    f_copy(base, a);
    for (int i = 0; i < 4; ++i) {
        uint64_t w = e[i];
        for (int b = 0; b < 64; ++b) {
            if (w & 1)
                f_mul(P, acc, acc, base);
            f_sqr(P, base, base);
            w >>= 1;
        }
    }
    f_copy(r, acc);
}

static void f_inv(const fparams* P, uint64_t* r, const uint64_t* a)
{
    /* Fermat: a^(m-2); 0 maps to 0 */
    uint64_t e[4];
    uint64_t two[4] = { 2, 0, 0, 0 };
    sub4(e, P->m, two);
    f_pow_limbs(P, r, a, e);
}

INL void f_from_mont(const fparams* P, uint64_t* r, const uint64_t* a)
{
    uint64_t one[4] = { 1, 0, 0, 0 };
    f_mul(P, r, a, one);
}

INL void f_to_mont(const fparams* P, uint64_t* r, const uint64_t* a) { f_mul(P, r, a, P->r2); }

static const fparams* fsel(int fld) { return fld == 0 ? &FR : &FQ; }

/* ---------------------------------------------------------------- threads */

typedef void (*range_fn)(void* ctx, size_t lo, size_t hi);
typedef struct {
    range_fn f;
    void* ctx;
    size_t lo, hi;
} job_t;

static void* job_run(void* arg)
{
    job_t* j = (job_t*)arg;
    j->f(j->ctx, j->lo, j->hi);
    return NULL;
}

int or_num_threads(void)
{
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (int)n : 1;
}

static void parallel_for(size_t n, int nthreads, range_fn f, void* ctx)
{
    if (nthreads <= 1 || n < 1024) {
        f(ctx, 0, n);
        return;
    }
    if ((size_t)nthreads > n)
        nthreads = (int)n;
    pthread_t th[256];
    job_t jobs[256];
    if (nthreads > 256)
        nthreads = 256;
    size_t chunk = (n + nthreads - 1) / nthreads;
    int started = 0;
    for (int t = 0; t < nthreads; ++t) {
        size_t lo = (size_t)t * chunk;
        size_t hi = lo + chunk > n ? n : lo + chunk;
        if (lo >= hi)
            break;
        jobs[t] = (job_t) { f, ctx, lo, hi };
        if (pthread_create(&th[t], NULL, job_run, &jobs[t]) != 0) {
            job_run(&jobs[t]);
            th[t] = 0;
        }
        started = t + 1;
    }
    for (int t = 0; t < started; ++t)
        if (th[t])
            pthread_join(th[t], NULL);
}

/* ------------------------------------------------------ field batch entry */

void or_field_mul(int fld, uint64_t* out, const uint64_t* a, const uint64_t* b, size_t n)
{
    const fparams* P = fsel(fld);
    for (size_t i = 0; i < n; ++i)
        f_mul(P, out + 4 * i, a + 4 * i, b + 4 * i);
}
void or_field_add(int fld, uint64_t* out, const uint64_t* a, const uint64_t* b, size_t n)
{
    const fparams* P = fsel(fld);
    for (size_t i = 0; i < n; ++i)
        f_add(P, out + 4 * i, a + 4 * i, b + 4 * i);
}
void or_field_sub(int fld, uint64_t* out, const uint64_t* a, const uint64_t* b, size_t n)
{
    const fparams* P = fsel(fld);
    for (size_t i = 0; i < n; ++i)
        f_sub(P, out + 4 * i, a + 4 * i, b + 4 * i);
}
void or_field_inv(int fld, uint64_t* out, const uint64_t* a, size_t n)
{
    const fparams* P = fsel(fld);
    for (size_t i = 0; i < n; ++i)
        f_inv(P, out + 4 * i, a + 4 * i);
}
void or_field_to_mont(int fld, uint64_t* out, const uint64_t* a, size_t n)
{
    const fparams* P = fsel(fld);
    for (size_t i = 0; i < n; ++i)
        f_to_mont(P, out + 4 * i, a + 4 * i);
}
void or_field_from_mont(int fld, uint64_t* out, const uint64_t* a, size_t n)
{
    const fparams* P = fsel(fld);
    for (size_t i = 0; i < n; ++i)
        f_from_mont(P, out + 4 * i, a + 4 * i);
}
void or_fr_pow(uint64_t* out, const uint64_t* base, uint64_t e)
{
    uint64_t ev[4] = { e, 0, 0, 0 };
    f_pow_limbs(&FR, out, base, ev);
}

/* halo2curves `from_u512`: d0 * R^2 + d1 * R^3 (Montgomery products), i.e.
 * the 512-bit little-endian value reduced mod r, in Montgomery form. */
static void fr_from_u512_one(uint64_t* out, const uint8_t* in)
{
    uint64_t d0[4], d1[4], r3[4], t0[4], t1[4];
    memcpy(d0, in, 32);
    memcpy(d1, in + 32, 32);
    f_mul(&FR, r3, FR.r2, FR.r2); /* R^4 / R = R^3 */
    /* d0 and d1 may exceed r: f_mul handles inputs < 2^256 because the CIOS
     * bound only needs a*b < m*2^256, true for a < 2^256 and b < m. */
    f_mul(&FR, t0, d0, FR.r2);
    f_mul(&FR, t1, d1, r3);
    f_add(&FR, out, t0, t1);
}

void or_fr_from_u512(uint64_t* out, const uint8_t* in, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        fr_from_u512_one(out + 4 * i, in + 64 * i);
}

/* ----------------------------------------------------------------- ChaCha */

#define ROTL32(v, c) (((v) << (c)) | ((v) >> (32 - (c))))
#define QR(a, b, c, d)                 \
    a += b;                            \
    d ^= a;                            \
    d = ROTL32(d, 16);                 \
    c += d;                            \
    b ^= c;                            \
    b = ROTL32(b, 12);                 \
    a += b;                            \
    d ^= a;                            \
    d = ROTL32(d, 8);                  \
    c += d;                            \
    b ^= c;                            \
    b = ROTL32(b, 7);

/* ChaCha20 block (RFC 8439 section 2.3) with a 64-bit block counter in words
 * 12..13 and a zero 64-bit stream id in words 14..15 (the rand_chacha layout;
 * identical to RFC 8439 with nonce 0 for counters below 2^32). */
void or_chacha_block(const uint32_t key[8], uint64_t counter, uint8_t out[64])
{
    uint32_t s[16] = { 0x61707865, 0x3320646e, 0x79622d32, 0x6b206574 };
    for (int i = 0; i < 8; ++i)
        s[4 + i] = key[i];
    s[12] = (uint32_t)counter;
    s[13] = (uint32_t)(counter >> 32);
    s[14] = 0;
    s[15] = 0;
    uint32_t x[16];
    memcpy(x, s, sizeof(x));
    for (int r = 0; r < 10; ++r) {
        QR(x[0], x[4], x[8], x[12]);
        QR(x[1], x[5], x[9], x[13]);
        QR(x[2], x[6], x[10], x[14]);
        QR(x[3], x[7], x[11], x[15]);
        QR(x[0], x[5], x[10], x[15]);
        QR(x[1], x[6], x[11], x[12]);
        QR(x[2], x[7], x[8], x[13]);
        QR(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; ++i) {
        uint32_t v = x[i] + s[i];
        out[4 * i + 0] = (uint8_t)v;
        out[4 * i + 1] = (uint8_t)(v >> 8);
        out[4 * i + 2] = (uint8_t)(v >> 16);
        out[4 * i + 3] = (uint8_t)(v >> 24);
    }
}

/* rand_core 0.6 `SeedableRng::seed_from_u64`: PCG32 (XSH-RR) words fill the seed. */
void or_chacha_seed_key(uint64_t seed, uint32_t key[8])
{
    const uint64_t MUL = 6364136223846793005ULL;
    const uint64_t INC = 11634580027462260723ULL;
    uint64_t state = seed;
    for (int i = 0; i < 8; ++i) {
        state = state * MUL + INC;
        uint32_t xorshifted = (uint32_t)(((state >> 18) ^ state) >> 27);
        uint32_t rot = (uint32_t)(state >> 59);
        key[i] = (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
    }
}

typedef struct {
    uint32_t key[8];
    uint64_t start;
    uint64_t* out;
} chacha_ctx;

static void chacha_range(void* vctx, size_t lo, size_t hi)
{
    chacha_ctx* c = (chacha_ctx*)vctx;
    uint8_t blk[64];
    for (size_t i = lo; i < hi; ++i) {
        or_chacha_block(c->key, c->start + i, blk);
        fr_from_u512_one(c->out + 4 * i, blk);
    }
}

void or_chacha_fr(uint64_t seed, uint64_t start, size_t n, uint64_t* out)
{
    chacha_ctx ctx;
    or_chacha_seed_key(seed, ctx.key);
    ctx.start = start;
    ctx.out = out;
    parallel_for(n, or_num_threads(), chacha_range, &ctx);
}

/* --------------------------------------------------------------------- G1 */
/* y^2 = x^3 + 3 over Fq, Jacobian (X, Y, Z); Z == 0 is the identity. */

typedef struct {
    uint64_t x[4], y[4], z[4];
} jac_t;

INL void j_set_inf(jac_t* r)
{
    f_zero(r->x);
    f_copy(r->y, FQ.one);
    f_zero(r->z);
}
INL int j_is_inf(const jac_t* p) { return f_is_zero(p->z); }
INL int aff_is_inf(const uint64_t* a) { return f_is_zero(a) && f_is_zero(a + 4); }

/* dbl-2009-l (a = 0) */
static void j_dbl(jac_t* r, const jac_t* p)
{
    const fparams* P = &FQ;
    if (j_is_inf(p)) {
        *r = *p;
        return;
    }
    uint64_t A[4], B[4], C[4], D[4], E[4], F[4], t[4];
    f_sqr(P, A, p->x);
    f_sqr(P, B, p->y);
    f_sqr(P, C, B);
    f_add(P, t, p->x, B);
    f_sqr(P, t, t);
    f_sub(P, t, t, A);
    f_sub(P, t, t, C);
    f_dbl(P, D, t);
    f_dbl(P, E, A);
    f_add(P, E, E, A);
    f_sqr(P, F, E);
    jac_t o;
    f_mul(P, o.z, p->y, p->z);
    f_dbl(P, o.z, o.z);
    f_dbl(P, t, D);
    f_sub(P, o.x, F, t);
    f_sub(P, t, D, o.x);
    f_mul(P, o.y, E, t);
    f_dbl(P, C, C);
    f_dbl(P, C, C);
    f_dbl(P, C, C);
    f_sub(P, o.y, o.y, C);
    *r = o;
}

/* add-2007-bl with the exceptional cases handled */
static void j_add(jac_t* r, const jac_t* p, const jac_t* q)
{
    const fparams* P = &FQ;
    if (j_is_inf(p)) {
        *r = *q;
        return;
    }
    if (j_is_inf(q)) {
        *r = *p;
        return;
    }
    uint64_t z1z1[4], z2z2[4], u1[4], u2[4], s1[4], s2[4], h[4], rr[4], t[4];
    f_sqr(P, z1z1, p->z);
    f_sqr(P, z2z2, q->z);
    f_mul(P, u1, p->x, z2z2);
    f_mul(P, u2, q->x, z1z1);
    f_mul(P, s1, p->y, q->z);
    f_mul(P, s1, s1, z2z2);
    f_mul(P, s2, q->y, p->z);
    f_mul(P, s2, s2, z1z1);
    f_sub(P, h, u2, u1);
    f_sub(P, rr, s2, s1);
    if (f_is_zero(h)) {
        if (f_is_zero(rr)) {
            j_dbl(r, p);
        } else {
            j_set_inf(r);
        }
        return;
    }
    f_dbl(P, rr, rr);
    uint64_t I[4], J[4], V[4];
    f_dbl(P, I, h);
    f_sqr(P, I, I);
    f_mul(P, J, h, I);
    f_mul(P, V, u1, I);
    jac_t o;
    f_sqr(P, o.x, rr);
    f_sub(P, o.x, o.x, J);
    f_dbl(P, t, V);
    f_sub(P, o.x, o.x, t);
    f_sub(P, t, V, o.x);
    f_mul(P, o.y, rr, t);
    f_mul(P, t, s1, J);
    f_dbl(P, t, t);
    f_sub(P, o.y, o.y, t);
    f_add(P, t, p->z, q->z);
    f_sqr(P, t, t);
    f_sub(P, t, t, z1z1);
    f_sub(P, t, t, z2z2);
    f_mul(P, o.z, t, h);
    *r = o;
}

/* madd-2007-bl: p Jacobian + q affine ((0,0) = identity) */
static void j_madd(jac_t* r, const jac_t* p, const uint64_t* q)
{
    const fparams* P = &FQ;
    if (aff_is_inf(q)) {
        *r = *p;
        return;
    }
    if (j_is_inf(p)) {
        f_copy(r->x, q);
        f_copy(r->y, q + 4);
        f_copy(r->z, FQ.one);
        return;
    }
    uint64_t z1z1[4], u2[4], s2[4], h[4], hh[4], I[4], J[4], rr[4], V[4], t[4];
    f_sqr(P, z1z1, p->z);
    f_mul(P, u2, q, z1z1);
    f_mul(P, s2, q + 4, p->z);
    f_mul(P, s2, s2, z1z1);
    f_sub(P, h, u2, p->x);
    f_sub(P, rr, s2, p->y);
    if (f_is_zero(h)) {
        if (f_is_zero(rr)) {
            j_dbl(r, p);
        } else {
            j_set_inf(r);
        }
        return;
    }
    f_dbl(P, rr, rr);
    f_sqr(P, hh, h);
    f_dbl(P, I, hh);
    f_dbl(P, I, I);
    f_mul(P, J, h, I);
    f_mul(P, V, p->x, I);
    jac_t o;
    f_sqr(P, o.x, rr);
    f_sub(P, o.x, o.x, J);
    f_dbl(P, t, V);
    f_sub(P, o.x, o.x, t);
    f_sub(P, t, V, o.x);
    f_mul(P, o.y, rr, t);
    f_mul(P, t, p->y, J);
    f_dbl(P, t, t);
    f_sub(P, o.y, o.y, t);
    f_add(P, t, p->z, h);
    f_sqr(P, t, t);
    f_sub(P, t, t, z1z1);
    f_sub(P, o.z, t, hh);
    *r = o;
}

static void j_to_affine(uint64_t* out_aff, const jac_t* p)
{
    if (j_is_inf(p)) {
        memset(out_aff, 0, 64);
        return;
    }
    uint64_t zi[4], zi2[4], zi3[4];
    f_inv(&FQ, zi, p->z);
    f_sqr(&FQ, zi2, zi);
    f_mul(&FQ, zi3, zi2, zi);
    f_mul(&FQ, out_aff, p->x, zi2);
    f_mul(&FQ, out_aff + 4, p->y, zi3);
}

/* the projective-out layout of the reference MSM: (x, y, 1) or (0, 1, 0) */
static void j_to_out12(uint64_t* out12, const jac_t* p)
{
    if (j_is_inf(p)) {
        memset(out12, 0, 96);
        f_copy(out12 + 4, FQ.one);
        return;
    }
    j_to_affine(out12, p);
    f_copy(out12 + 8, FQ.one);
}

void or_g1_generator(uint64_t* out_affine)
{
    uint64_t one[4] = { 1, 0, 0, 0 }, two[4] = { 2, 0, 0, 0 };
    f_to_mont(&FQ, out_affine, one);
    f_to_mont(&FQ, out_affine + 4, two);
}

int or_g1_is_on_curve(const uint64_t* a)
{
    if (aff_is_inf(a))
        return 1;
    uint64_t three[4] = { 3, 0, 0, 0 }, b[4], y2[4], x3[4];
    f_to_mont(&FQ, b, three);
    f_sqr(&FQ, y2, a + 4);
    f_sqr(&FQ, x3, a);
    f_mul(&FQ, x3, x3, a);
    f_add(&FQ, x3, x3, b);
    return f_eq(x3, y2);
}

static void j_mul_affine(jac_t* r, const uint64_t* p_aff, const uint64_t* s_mont)
{
    uint64_t s[4];
    f_from_mont(&FR, s, s_mont);
    jac_t acc;
    j_set_inf(&acc);
    for (int i = 255; i >= 0; --i) {
        j_dbl(&acc, &acc);
        if ((s[i / 64] >> (i % 64)) & 1)
            j_madd(&acc, &acc, p_aff);
    }
    *r = acc;
}

void or_g1_mul(uint64_t* out12, const uint64_t* p_affine, const uint64_t* s)
{
    jac_t r;
    j_mul_affine(&r, p_affine, s);
    j_to_out12(out12, &r);
}

void or_g1_add_affine(uint64_t* out12, const uint64_t* p, const uint64_t* q)
{
    jac_t a;
    j_set_inf(&a);
    j_madd(&a, &a, p);
    j_madd(&a, &a, q);
    j_to_out12(out12, &a);
}

typedef struct {
    uint64_t* out;
    const uint64_t* scalars;
    const uint64_t* table; /* 2^i * G affine, i < 254 */
} mulgen_ctx;

static void mulgen_range(void* vctx, size_t lo, size_t hi)
{
    mulgen_ctx* c = (mulgen_ctx*)vctx;
    for (size_t i = lo; i < hi; ++i) {
        uint64_t s[4];
        f_from_mont(&FR, s, c->scalars + 4 * i);
        jac_t acc;
        j_set_inf(&acc);
        for (int b = 0; b < 254; ++b)
            if ((s[b / 64] >> (b % 64)) & 1)
                j_madd(&acc, &acc, c->table + 8 * b);
        j_to_affine(c->out + 8 * i, &acc);
    }
}

void or_g1_mul_gen(uint64_t* out_affine, const uint64_t* scalars, size_t n, int nthreads)
{
    uint64_t* table = (uint64_t*)malloc(254 * 64);
    jac_t g;
    uint64_t gen[8];
    or_g1_generator(gen);
    j_set_inf(&g);
    j_madd(&g, &g, gen);
    for (int b = 0; b < 254; ++b) {
        j_to_affine(table + 8 * b, &g);
        j_dbl(&g, &g);
    }
    mulgen_ctx ctx = { out_affine, scalars, table };
    parallel_for(n, nthreads, mulgen_range, &ctx);
    free(table);
}

/* -------------------------------------------------------------------- MSM */
/* Restates src/cpu/arithmetic.rs:14-129 (multiexp_serial): unsigned c-bit
 * windows over the canonical little-endian repr, c = ceil(ln n) (1 below 4
 * points, 3 below 32), buckets holding None / Affine / Projective, running-sum
 * bucket reduction, and c doublings of the accumulator per window. */

static size_t get_at(size_t segment, size_t c, const uint8_t* bytes)
{
    size_t skip_bits = segment * c;
    size_t skip_bytes = skip_bits / 8;
    if (skip_bytes >= 32)
        return 0;
    uint8_t v[8] = { 0 };
    for (size_t k = 0; k < 8 && skip_bytes + k < 32; ++k)
        v[k] = bytes[skip_bytes + k];
    uint64_t tmp;
    memcpy(&tmp, v, 8);
    tmp >>= skip_bits - skip_bytes * 8;
    tmp %= ((uint64_t)1 << c);
    return (size_t)tmp;
}

enum { B_NONE = 0, B_AFF = 1, B_PROJ = 2 };
typedef struct {
    int kind;
    const uint64_t* aff;
    jac_t proj;
} bucket_t;

static void multiexp_serial(const uint64_t* coeffs_mont, const uint64_t* bases, size_t n, jac_t* acc)
{
    size_t c = n < 4 ? 1 : (n < 32 ? 3 : (size_t)ceil(log((double)(uint32_t)n)));
    uint8_t* repr = (uint8_t*)malloc(n * 32 + 1);
    for (size_t i = 0; i < n; ++i) {
        uint64_t t[4];
        f_from_mont(&FR, t, coeffs_mont + 4 * i);
        memcpy(repr + 32 * i, t, 32);
    }
    size_t segments = (254 + c - 1) / c;
    long max_nonzero = -1;
    for (size_t s = 0; s < segments; ++s)
        for (size_t i = 0; i < n; ++i)
            if (get_at(s, c, repr + 32 * i) != 0) {
                max_nonzero = (long)s;
                break;
            }
    if (max_nonzero < 0) {
        free(repr);
        return;
    }
    size_t nb = ((size_t)1 << c) - 1;
    bucket_t* buckets = (bucket_t*)malloc(sizeof(bucket_t) * (nb ? nb : 1));
    for (long s = max_nonzero; s >= 0; --s) {
        for (size_t k = 0; k < c; ++k)
            j_dbl(acc, acc);
        for (size_t b = 0; b < nb; ++b)
            buckets[b].kind = B_NONE;
        size_t max_bits = 0;
        for (size_t i = 0; i < n; ++i) {
            size_t d = get_at((size_t)s, c, repr + 32 * i);
            if (d == 0)
                continue;
            if (d > max_bits)
                max_bits = d;
            bucket_t* bk = &buckets[d - 1];
            const uint64_t* base = bases + 8 * i;
            if (bk->kind == B_NONE) {
                bk->kind = B_AFF;
                bk->aff = base;
            } else if (bk->kind == B_AFF) {
                jac_t t;
                j_set_inf(&t);
                j_madd(&t, &t, bk->aff);
                j_madd(&bk->proj, &t, base);
                bk->kind = B_PROJ;
            } else {
                j_madd(&bk->proj, &bk->proj, base);
            }
        }
        jac_t running;
        j_set_inf(&running);
        for (size_t b = max_bits; b-- > 0;) {
            if (buckets[b].kind == B_AFF)
                j_madd(&running, &running, buckets[b].aff);
            else if (buckets[b].kind == B_PROJ)
                j_add(&running, &running, &buckets[b].proj);
            j_add(acc, acc, &running);
        }
    }
    free(buckets);
    free(repr);
}

typedef struct {
    const uint64_t* coeffs;
    const uint64_t* bases;
    size_t n, chunk;
    jac_t* results;
} msm_ctx;

static void msm_range(void* vctx, size_t lo, size_t hi)
{
    msm_ctx* c = (msm_ctx*)vctx;
    for (size_t k = lo; k < hi; ++k) {
        size_t off = k * c->chunk;
        size_t len = off + c->chunk > c->n ? c->n - off : c->chunk;
        j_set_inf(&c->results[k]);
        multiexp_serial(c->coeffs + 4 * off, c->bases + 8 * off, len, &c->results[k]);
    }
}

static void thread_each(size_t count, range_fn f, void* ctx)
{
    /* one thread per item (best_multiexp_cpu spawns one task per chunk) */
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * count);
    job_t* jobs = (job_t*)malloc(sizeof(job_t) * count);
    for (size_t k = 0; k < count; ++k) {
        jobs[k] = (job_t) { f, ctx, k, k + 1 };
        if (pthread_create(&th[k], NULL, job_run, &jobs[k]) != 0) {
            job_run(&jobs[k]);
            th[k] = 0;
        }
    }
    for (size_t k = 0; k < count; ++k)
        if (th[k])
            pthread_join(th[k], NULL);
    free(th);
    free(jobs);
}

/* src/cpu/arithmetic.rs:131-156 (best_multiexp_cpu) */
void or_msm(uint64_t* out12, const uint64_t* scalars, const uint64_t* points, size_t n, int nthreads)
{
    if (nthreads < 1)
        nthreads = 1;
    jac_t acc;
    j_set_inf(&acc);
    if (n > (size_t)nthreads && nthreads > 1) {
        size_t chunk = n / (size_t)nthreads;
        size_t nchunks = (n + chunk - 1) / chunk;
        jac_t* res = (jac_t*)malloc(sizeof(jac_t) * nchunks);
        msm_ctx ctx = { scalars, points, n, chunk, res };
        thread_each(nchunks, msm_range, &ctx);
        for (size_t k = 0; k < nchunks; ++k)
            j_add(&acc, &acc, &res[k]);
        free(res);
    } else {
        multiexp_serial(scalars, points, n, &acc);
    }
    j_to_out12(out12, &acc);
}

void or_msm_naive(uint64_t* out12, const uint64_t* scalars, const uint64_t* points, size_t n)
{
    jac_t acc, t;
    j_set_inf(&acc);
    for (size_t i = 0; i < n; ++i) {
        j_mul_affine(&t, points + 8 * i, scalars + 4 * i);
        j_add(&acc, &acc, &t);
    }
    j_to_out12(out12, &acc);
}

/* -------------------------------------------------------------------- NTT */

typedef struct {
    uint64_t* a;
    const uint64_t* tw;
    size_t n, half, tw_stride;
    uint32_t log_n;
    const uint64_t* c;
} ntt_ctx;

static uint32_t bitrev(uint32_t x, uint32_t bits)
{
    uint32_t r = 0;
    for (uint32_t i = 0; i < bits; ++i) {
        r = (r << 1) | (x & 1);
        x >>= 1;
    }
    return r;
}

static void revbin_range(void* vctx, size_t lo, size_t hi)
{
    ntt_ctx* c = (ntt_ctx*)vctx;
    for (size_t i = lo; i < hi; ++i) {
        size_t j = bitrev((uint32_t)i, c->log_n);
        if (i < j) {
            uint64_t t[4];
            f_copy(t, c->a + 4 * i);
            f_copy(c->a + 4 * i, c->a + 4 * j);
            f_copy(c->a + 4 * j, t);
        }
    }
}

static void stage_range(void* vctx, size_t lo, size_t hi)
{
    ntt_ctx* c = (ntt_ctx*)vctx;
    size_t m = c->half;
    for (size_t b = lo; b < hi; ++b) {
        size_t k = b / m, j = b % m;
        uint64_t* x = c->a + 4 * (k * 2 * m + j);
        uint64_t* y = x + 4 * m;
        uint64_t t[4], u[4];
        f_mul(&FR, t, y, c->tw + 4 * (j * c->tw_stride));
        f_copy(u, x);
        f_add(&FR, x, u, t);
        f_sub(&FR, y, u, t);
    }
}

typedef struct {
    uint64_t* tw;
    const uint64_t* omega;
} tw_ctx;

static void tw_range(void* vctx, size_t lo, size_t hi)
{
    tw_ctx* c = (tw_ctx*)vctx;
    uint64_t cur[4];
    or_fr_pow(cur, c->omega, (uint64_t)lo);
    for (size_t i = lo; i < hi; ++i) {
        f_copy(c->tw + 4 * i, cur);
        f_mul(&FR, cur, cur, c->omega);
    }
}

/* Radix-2 decimation-in-time DFT: out[j] = sum_i a[i] omega^(i j), natural
 * order in and out (the semantics of halo2's best_fft, src/fft.rs:16-24). */
void or_fft(uint64_t* a, uint32_t log_n, const uint64_t* omega, int nthreads)
{
    size_t n = (size_t)1 << log_n;
    if (n == 1)
        return;
    uint64_t* tw = (uint64_t*)malloc(32 * (n / 2));
    tw_ctx tctx = { tw, omega };
    parallel_for(n / 2, nthreads, tw_range, &tctx);
    ntt_ctx ctx = { a, tw, n, 0, 0, log_n, NULL };
    parallel_for(n, nthreads, revbin_range, &ctx);
    for (size_t m = 1; m < n; m <<= 1) {
        ctx.half = m;
        ctx.tw_stride = n / (2 * m);
        parallel_for(n / 2, nthreads, stage_range, &ctx);
    }
    free(tw);
}

void or_dft_naive(uint64_t* out, const uint64_t* a, uint32_t log_n, const uint64_t* omega)
{
    size_t n = (size_t)1 << log_n;
    uint64_t wj[4], w[4], acc[4], t[4];
    f_copy(wj, FR.one);
    for (size_t j = 0; j < n; ++j) {
        f_zero(acc);
        f_copy(w, FR.one);
        for (size_t i = 0; i < n; ++i) {
            f_mul(&FR, t, a + 4 * i, w);
            f_add(&FR, acc, acc, t);
            f_mul(&FR, w, w, wj);
        }
        f_copy(out + 4 * j, acc);
        f_mul(&FR, wj, wj, omega);
    }
}

static void scale_range(void* vctx, size_t lo, size_t hi)
{
    ntt_ctx* c = (ntt_ctx*)vctx;
    for (size_t i = lo; i < hi; ++i)
        f_mul(&FR, c->a + 4 * i, c->a + 4 * i, c->c);
}

void or_scale(uint64_t* a, size_t n, const uint64_t* cst, int nthreads)
{
    ntt_ctx ctx = { a, NULL, n, 0, 0, 0, cst };
    parallel_for(n, nthreads, scale_range, &ctx);
}

static void powers_range(void* vctx, size_t lo, size_t hi)
{
    ntt_ctx* c = (ntt_ctx*)vctx;
    uint64_t cur[4];
    or_fr_pow(cur, c->c, (uint64_t)lo);
    for (size_t i = lo; i < hi; ++i) {
        f_mul(&FR, c->a + 4 * i, c->a + 4 * i, cur);
        f_mul(&FR, cur, cur, c->c);
    }
}

/* a[i] *= c^i (src/cpu/evaluator.rs:26-45, the CosetFFT_Part twist) */
void or_distribute_powers(uint64_t* a, size_t n, const uint64_t* cst, int nthreads)
{
    ntt_ctx ctx = { a, NULL, n, 0, 0, 0, cst };
    parallel_for(n, nthreads, powers_range, &ctx);
}

/* cuda/include/kernel/polynomial.h:58-75 */
void or_distribute_powers_zeta(uint64_t* a, size_t n, const uint64_t* coset_powers, uint32_t len)
{
    uint32_t mod = len + 1;
    for (size_t i = 0; i < n; ++i) {
        uint32_t k = (uint32_t)(i % mod);
        if (k != 0)
            f_mul(&FR, a + 4 * i, a + 4 * i, coset_powers + 4 * (k - 1));
    }
}

/* cuda/include/kernel/polynomial.h:35-53 */
void or_divide_by_vanishing(uint64_t* a, size_t n, const uint64_t* t_evals, uint32_t t_len)
{
    for (size_t i = 0; i < n; ++i)
        f_mul(&FR, a + 4 * i, a + 4 * i, t_evals + 4 * (i % t_len));
}

/* cuda/include/kernel/polynomial.h:16-30 */
void or_extended_from_lagrange_vec(uint64_t* out, const uint64_t* const* parts, uint32_t num_parts, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        for (uint32_t p = 0; p < num_parts; ++p)
            f_copy(out + 4 * (i * num_parts + p), parts[p] + 4 * i);
}

/* -------------------------------------------------------------- products */

/* Montgomery's trick; zeros stay zero (ff::BatchInvert semantics). */
void or_batch_invert(int fld, uint64_t* a, size_t n)
{
    const fparams* P = fsel(fld);
    if (n == 0)
        return;
    uint64_t* pre = (uint64_t*)malloc(32 * n);
    uint64_t acc[4], t[4];
    f_copy(acc, P->one);
    for (size_t i = 0; i < n; ++i) {
        f_copy(pre + 4 * i, acc);
        if (!f_is_zero(a + 4 * i))
            f_mul(P, acc, acc, a + 4 * i);
    }
    f_inv(P, acc, acc);
    for (size_t i = n; i-- > 0;) {
        if (f_is_zero(a + 4 * i))
            continue;
        f_mul(P, t, acc, pre + 4 * i);
        f_mul(P, acc, acc, a + 4 * i);
        f_copy(a + 4 * i, t);
    }
    free(pre);
}

/* out[i] = prefix * prod_{j <= i} in[j] (src/cuda/funcs/grand_product.rs:14-74,
 * src/cuda/tests.rs:1430-1436) */
void or_grand_product(uint64_t* out, const uint64_t* in, const uint64_t* prefix, size_t n)
{
    uint64_t acc[4];
    f_copy(acc, prefix);
    for (size_t i = 0; i < n; ++i) {
        f_mul(&FR, acc, acc, in + 4 * i);
        f_copy(out + 4 * i, acc);
    }
}

/* src/cpu/arithmetic.rs:158-197 (lookup_product_cpu) */
void or_lookup_product(uint64_t* out, const uint64_t* pi, const uint64_t* pt, const uint64_t* ci,
                       const uint64_t* ct, const uint64_t* beta, const uint64_t* gamma, size_t n)
{
    uint64_t a[4], b[4];
    for (size_t i = 0; i < n; ++i) {
        f_add(&FR, a, beta, pi + 4 * i);
        f_add(&FR, b, gamma, pt + 4 * i);
        f_mul(&FR, out + 4 * i, a, b);
    }
    or_batch_invert(0, out, n);
    for (size_t i = 0; i < n; ++i) {
        f_add(&FR, a, ci + 4 * i, beta);
        f_mul(&FR, out + 4 * i, out + 4 * i, a);
        f_add(&FR, b, ct + 4 * i, gamma);
        f_mul(&FR, out + 4 * i, out + 4 * i, b);
    }
}

/* src/cpu/arithmetic.rs:287-327 (permutation_product_cpu); deltaomega is
 * advanced by delta once per column, as the device entry point does. */
void or_permutation_product(uint64_t* modified, const uint64_t* const* perms, const uint64_t* const* values,
                            size_t batch, const uint64_t* beta, const uint64_t* gamma, const uint64_t* delta,
                            const uint64_t* omega, uint64_t* deltaomega, size_t n)
{
    uint64_t t[4];
    for (size_t c = 0; c < batch; ++c)
        for (size_t i = 0; i < n; ++i) {
            f_mul(&FR, t, beta, perms[c] + 4 * i);
            f_add(&FR, t, t, gamma);
            f_add(&FR, t, t, values[c] + 4 * i);
            f_mul(&FR, modified + 4 * i, modified + 4 * i, t);
        }
    or_batch_invert(0, modified, n);
    for (size_t c = 0; c < batch; ++c) {
        uint64_t dw[4];
        f_copy(dw, deltaomega);
        for (size_t i = 0; i < n; ++i) {
            f_mul(&FR, t, dw, beta);
            f_add(&FR, t, t, gamma);
            f_add(&FR, t, t, values[c] + 4 * i);
            f_mul(&FR, modified + 4 * i, modified + 4 * i, t);
            f_mul(&FR, dw, dw, omega);
        }
        f_mul(&FR, deltaomega, deltaomega, delta);
    }
}

static size_t rot_idx(size_t idx, int32_t rot, int32_t rot_scale, int32_t isize_)
{
    int64_t v = (int64_t)idx + (int64_t)rot * rot_scale;
    int64_t m = v % isize_;
    if (m < 0)
        m += isize_;
    return (size_t)m;
}

/* src/cpu/arithmetic.rs:223-280 (quotient_lookups_cpu), isize = n */
void or_quotient_lookups(uint64_t* values, const uint64_t* table_values, const uint64_t* z,
                         const uint64_t* ai, const uint64_t* st, const uint64_t* l0, const uint64_t* l_last,
                         const uint64_t* l_active, const uint64_t* beta, const uint64_t* gamma,
                         const uint64_t* y, size_t n)
{
    uint64_t v[4], t[4], u[4], ams[4];
    for (size_t idx = 0; idx < n; ++idx) {
        size_t rn = rot_idx(idx, 1, 1, (int32_t)n), rp = rot_idx(idx, -1, 1, (int32_t)n);
        f_copy(v, values + 4 * idx);
        f_sub(&FR, ams, ai + 4 * idx, st + 4 * idx);
        /* l_0 (1 - z) */
        f_sub(&FR, t, FR.one, z + 4 * idx);
        f_mul(&FR, t, t, l0 + 4 * idx);
        f_mul(&FR, v, v, y);
        f_add(&FR, v, v, t);
        /* l_last (z^2 - z) */
        f_sqr(&FR, t, z + 4 * idx);
        f_sub(&FR, t, t, z + 4 * idx);
        f_mul(&FR, t, t, l_last + 4 * idx);
        f_mul(&FR, v, v, y);
        f_add(&FR, v, v, t);
        /* l_active (z(wX)(a'+b)(s'+g) - z * table) */
        f_add(&FR, t, ai + 4 * idx, beta);
        f_mul(&FR, t, z + 4 * rn, t);
        f_add(&FR, u, st + 4 * idx, gamma);
        f_mul(&FR, t, t, u);
        f_mul(&FR, u, z + 4 * idx, table_values + 4 * idx);
        f_sub(&FR, t, t, u);
        f_mul(&FR, t, t, l_active + 4 * idx);
        f_mul(&FR, v, v, y);
        f_add(&FR, v, v, t);
        /* l_0 (a' - s') */
        f_mul(&FR, t, ams, l0 + 4 * idx);
        f_mul(&FR, v, v, y);
        f_add(&FR, v, v, t);
        /* l_active (a' - s')(a' - a'(w^-1 X)) */
        f_sub(&FR, t, ai + 4 * idx, ai + 4 * rp);
        f_mul(&FR, t, ams, t);
        f_mul(&FR, t, t, l_active + 4 * idx);
        f_mul(&FR, v, v, y);
        f_add(&FR, v, v, t);
        f_copy(values + 4 * idx, v);
    }
}

/* src/cpu/evaluator.rs:63-147 (permutation_quotient_cpu_chunk), start = 0 */
void or_permutation_quotient(uint64_t* values, const uint64_t* const* zs, const uint64_t* const* perms,
                             const uint64_t* const* cols, const uint64_t* l0, const uint64_t* l_last,
                             const uint64_t* l_active, const uint64_t* beta, const uint64_t* gamma,
                             const uint64_t* y, const uint64_t* delta, const uint64_t* delta_start,
                             const uint64_t* cur_ext_omega, const uint64_t* omega, size_t n_sets,
                             size_t chunk_len, size_t n_perm_cols, int32_t last_rotation, int32_t rot_scale,
                             int32_t isize_, size_t n)
{
    if (n_sets == 0)
        return;
    uint64_t beta_term[4], v[4], t[4], cd[4], left[4], right[4], sc[4];
    f_copy(beta_term, cur_ext_omega);
    for (size_t idx = 0; idx < n; ++idx) {
        size_t rn = rot_idx(idx, 1, rot_scale, isize_);
        size_t rl = rot_idx(idx, last_rotation, rot_scale, isize_);
        f_copy(v, values + 4 * idx);
        f_sub(&FR, t, FR.one, zs[0] + 4 * idx);
        f_mul(&FR, t, t, l0 + 4 * idx);
        f_mul(&FR, v, v, y);
        f_add(&FR, v, v, t);
        const uint64_t* zl = zs[n_sets - 1] + 4 * idx;
        f_sqr(&FR, t, zl);
        f_sub(&FR, t, t, zl);
        f_mul(&FR, t, t, l_last + 4 * idx);
        f_mul(&FR, v, v, y);
        f_add(&FR, v, v, t);
        for (size_t s = 1; s < n_sets; ++s) {
            f_sub(&FR, t, zs[s] + 4 * idx, zs[s - 1] + 4 * rl);
            f_mul(&FR, t, t, l0 + 4 * idx);
            f_mul(&FR, v, v, y);
            f_add(&FR, v, v, t);
        }
        f_mul(&FR, cd, delta_start, beta_term);
        for (size_t s = 0; s < n_sets; ++s) {
            f_copy(left, zs[s] + 4 * rn);
            f_copy(right, zs[s] + 4 * idx);
            size_t c0 = s * chunk_len;
            size_t cnt = c0 + chunk_len <= n_perm_cols ? chunk_len : n_perm_cols - c0;
            for (size_t j = 0; j < cnt; ++j) {
                const uint64_t* cv = cols[c0 + j] + 4 * idx;
                f_add(&FR, sc, cv, gamma);
                f_mul(&FR, t, beta, perms[c0 + j] + 4 * idx);
                f_add(&FR, t, t, sc);
                f_mul(&FR, left, left, t);
                f_add(&FR, t, sc, cd);
                f_mul(&FR, right, right, t);
                f_mul(&FR, cd, cd, delta);
            }
            f_sub(&FR, t, left, right);
            f_mul(&FR, t, t, l_active + 4 * idx);
            f_mul(&FR, v, v, y);
            f_add(&FR, v, v, t);
        }
        f_copy(values + 4 * idx, v);
        f_mul(&FR, beta_term, beta_term, omega);
    }
}

/* sum_i a[i] * b[i] in Fr (closed-form MSM checks: sum c_i (s_i G) = (sum c_i s_i) G) */
typedef struct {
    const uint64_t *a, *b;
    uint64_t part[256][4];
    size_t chunk;
    size_t n;
} ip_ctx;

static void ip_range(void* vctx, size_t lo, size_t hi)
{
    ip_ctx* c = (ip_ctx*)vctx;
    uint64_t acc[4] = { 0, 0, 0, 0 }, t[4];
    for (size_t i = lo; i < hi; ++i) {
        f_mul(&FR, t, c->a + 4 * i, c->b + 4 * i);
        f_add(&FR, acc, acc, t);
    }
    f_copy(c->part[lo / c->chunk], acc);
}

void or_fr_inner_product(uint64_t* out, const uint64_t* a, const uint64_t* b, size_t n)
{
    ip_ctx* ctx = (ip_ctx*)calloc(1, sizeof(ip_ctx));
    int nt = or_num_threads();
    if (nt > 256)
        nt = 256;
    ctx->a = a;
    ctx->b = b;
    ctx->n = n;
    ctx->chunk = (n + nt - 1) / nt;
    if (ctx->chunk == 0)
        ctx->chunk = 1;
    if (n < 1024) {
        ip_range(ctx, 0, n);
        f_copy(out, ctx->part[0]);
        free(ctx);
        return;
    }
    parallel_for(n, nt, ip_range, ctx);
    uint64_t acc[4] = { 0, 0, 0, 0 };
    for (int t = 0; t < 256; ++t)
        f_add(&FR, acc, acc, ctx->part[t]);
    f_copy(out, acc);
    free(ctx);
}

/* Custom-gate rule interpreter, restating the row semantics of
 * cuda/include/kernel/quotient.h:57-303 (equivalently Calculation::evaluate,
 * src/plonk/evaluation.rs:215-266, folded with Horner in y as
 * src/cpu/evaluator.rs:150-411 does): intermediate i = (C1 a + C2 b)(d ? b : 1),
 * C = [0, 1, -1, 2]; out[r] = Horner_y(prev[r] or 0, value parts). */
void or_quotient_rules(uint64_t* out, const uint64_t* prev, const uint64_t* const* fixed, size_t nf,
                       const uint64_t* const* instance, size_t ni, const uint64_t* const* advice, size_t na,
                       const uint64_t* challenges, const uint64_t* constants, const uint64_t* y,
                       const uint64_t* rules, size_t num_rules, const uint64_t* vp, size_t num_vp, size_t nq,
                       int64_t rot_scale, size_t row_start, size_t row_end)
{
    (void)nf;
    (void)ni;
    (void)na;
    uint64_t* interm = (uint64_t*)calloc(num_rules ? num_rules : 1, 32);
    for (size_t r = row_start; r < row_end; ++r) {
        for (size_t i = 0; i < num_rules; ++i) {
            uint64_t w[2] = { rules[2 * i], rules[2 * i + 1] };
            uint64_t v[2][4];
            for (int o = 0; o < 2; ++o) {
                uint32_t src = (uint32_t)(w[o] & 0xf), idx = (uint32_t)((w[o] >> 4) & 0xfffff);
                int64_t rot = (int64_t)((w[o] >> 24) & 0x7fff);
                if ((w[o] >> 39) & 1)
                    rot = -rot;
                int64_t rr = ((int64_t)r + rot * rot_scale) % (int64_t)nq;
                if (rr < 0)
                    rr += (int64_t)nq;
                const uint64_t* s;
                static const uint64_t zero4[4] = { 0, 0, 0, 0 };
                switch (src) {
                case 0: s = fixed[idx] + 4 * rr; break;
                case 1: s = instance[idx] + 4 * rr; break;
                case 2: s = advice[idx] + 4 * rr; break;
                case 3: s = interm + 4 * idx; break;
                case 4: s = constants + 4 * idx; break;
                case 5: s = challenges + 4 * idx; break;
                default: s = zero4; break;
                }
                f_copy(v[o], s);
            }
            uint32_t c1 = (uint32_t)((w[0] >> 40) & 0xf), d = (uint32_t)((w[0] >> 44) & 0xf);
            uint32_t c2 = (uint32_t)((w[1] >> 40) & 0xf);
            uint64_t res[4] = { 0, 0, 0, 0 };
            if (c1 == 1) {
                if (c2 == 0) {
                    if (d == 1)
                        f_mul(&FR, res, v[0], v[1]);
                    else
                        f_copy(res, v[0]);
                } else if (c2 == 1) {
                    f_add(&FR, res, v[0], v[1]);
                } else if (c2 == 2) {
                    f_sub(&FR, res, v[0], v[1]);
                }
            } else if (c1 == 2) {
                uint64_t z[4] = { 0, 0, 0, 0 };
                f_sub(&FR, res, z, v[0]);
            } else if (c1 == 3) {
                f_add(&FR, res, v[0], v[0]);
            }
            f_copy(interm + 4 * i, res);
        }
        uint64_t val[4] = { 0, 0, 0, 0 };
        if (prev)
            f_copy(val, prev + 4 * (r - row_start));
        for (size_t i = 0; i < num_vp; ++i) {
            uint32_t src = (uint32_t)(vp[i] & 0xf), idx = (uint32_t)((vp[i] >> 4) & 0xfffff);
            const uint64_t* part = src == 4 ? constants + 4 * idx : interm + 4 * idx;
            f_mul(&FR, val, val, y);
            f_add(&FR, val, val, part);
        }
        f_copy(out + 4 * (r - row_start), val);
    }
    free(interm);
}

/* a(u) = sum a_i u^i by Horner from the top coefficient: halo2-axiom 0.5.3
 * (git 5e4f0e52) arithmetic::eval_polynomial, the CPU side of the reference's
 * src/cuda/tests.rs:53-68 and tests/eval_polynomial_horner_device_equivalence.rs. */
void or_eval_poly(uint64_t* out, const uint64_t* a, size_t n, const uint64_t* u)
{
    uint64_t acc[4] = { 0, 0, 0, 0 };
    for (size_t i = n; i-- > 0;) {
        f_mul(&FR, acc, acc, u);
        f_add(&FR, acc, acc, a + 4 * i);
    }
    f_copy(out, acc);
}

/* q = (a(X) - a(u)) / (X - u), n - 1 coefficients: halo2-axiom 0.5.3
 * arithmetic::kate_division (re-exported at src/cpu/arithmetic.rs:11; compared
 * with the device path in src/cuda/tests.rs:2830-2896). Walking the coefficients
 * from the top: lead = a_i - tmp, q_{i-1} = lead, tmp = lead * (-u). */
void or_kate_division(uint64_t* q, const uint64_t* a, size_t n, const uint64_t* u)
{
    uint64_t neg_u[4], tmp[4] = { 0, 0, 0, 0 }, lead[4];
    const uint64_t zero[4] = { 0, 0, 0, 0 };
    f_sub(&FR, neg_u, zero, u);
    for (size_t i = n; i-- > 1;) {
        f_sub(&FR, lead, a + 4 * i, tmp);
        f_copy(q + 4 * (i - 1), lead);
        f_mul(&FR, tmp, lead, neg_u);
    }
}

/* ---- lookup permutation pair (src/plonk/lookup/prover.rs:788-882, permute_expression_pair_seq) ---- */

typedef struct {
    uint64_t c[4]; /* canonical value */
    uint64_t m[4]; /* Montgomery form as stored */
} perm_key;

static int perm_cmp(const void* pa, const void* pb)
{
    const perm_key* a = (const perm_key*)pa;
    const perm_key* b = (const perm_key*)pb;
    for (int i = 3; i >= 0; --i)
        if (a->c[i] != b->c[i])
            return a->c[i] < b->c[i] ? -1 : 1;
    return 0;
}

/* permuted_input = the first U inputs sorted by canonical value; at the first row
 * of each run of equal inputs permuted_table takes that value and one table
 * occurrence of it is consumed; the leftover table values, ascending, fill the
 * repeated rows from the last one backwards; rows [U, n) are zero. Returns 1
 * (outputs unspecified) when an input value is missing from the table, where
 * the reference panics. */
int or_permute_expression_pair(uint64_t* pin, uint64_t* ptab, const uint64_t* input, const uint64_t* table,
                               size_t n, size_t U)
{
    memset(pin, 0, 32 * n);
    memset(ptab, 0, 32 * n);
    if (U == 0)
        return 0;
    perm_key* in = (perm_key*)malloc(U * sizeof(perm_key));
    perm_key* tab = (perm_key*)malloc(U * sizeof(perm_key));
    uint32_t* cnt = (uint32_t*)malloc(U * sizeof(uint32_t));
    size_t* rep = (size_t*)malloc(U * sizeof(size_t));
    for (size_t i = 0; i < U; ++i) {
        f_copy(in[i].m, input + 4 * i);
        f_from_mont(&FR, in[i].c, input + 4 * i);
        f_copy(tab[i].m, table + 4 * i);
        f_from_mont(&FR, tab[i].c, table + 4 * i);
    }
    qsort(in, U, sizeof(perm_key), perm_cmp);
    qsort(tab, U, sizeof(perm_key), perm_cmp);
    /* distinct table values (ascending) with their counts, compacted to the front */
    size_t nd = 0;
    for (size_t j = 0; j < U; ++j) {
        if (nd > 0 && perm_cmp(&tab[nd - 1], &tab[j]) == 0) {
            ++cnt[nd - 1];
        } else {
            tab[nd] = tab[j];
            cnt[nd] = 1;
            ++nd;
        }
    }
    int err = 0;
    size_t nrep = 0;
    for (size_t row = 0; row < U; ++row) {
        f_copy(pin + 4 * row, in[row].m);
        if (row == 0 || perm_cmp(&in[row], &in[row - 1]) != 0) {
            f_copy(ptab + 4 * row, in[row].m);
            perm_key* hit = (perm_key*)bsearch(&in[row], tab, nd, sizeof(perm_key), perm_cmp);
            if (!hit) {
                err = 1;
                break;
            }
            --cnt[hit - tab];
        } else {
            rep[nrep++] = row;
        }
    }
    for (size_t d = 0; d < nd && !err; ++d)
        for (uint32_t c = 0; c < cnt[d]; ++c) {
            if (nrep == 0) {
                err = 1;
                break;
            }
            f_copy(ptab + 4 * rep[--nrep], tab[d].m);
        }
    if (nrep)
        err = 1;
    free(in);
    free(tab);
    free(cnt);
    free(rep);
    return err;
}
