/*
 * CPU oracle for the halo2-axiom-gpu hot path: a plain-C restatement of the
 * reference CPU algorithms, used ONLY as the checker in tests, the smoke test
 * and the CPU-baseline leg of the benchmark. Nothing in the product links it.
 *
 * Representation (identical to the device boundary):
 *   Fr / Fq element : 4 x uint64 little-endian limbs, Montgomery form a*2^256 mod m,
 *                     fully reduced (< m).
 *   G1 affine       : (x, y) Fq, 8 limbs; the identity is (0, 0).
 *   G1 "projective" : (X, Y, Z) Fq, 12 limbs, written as (x, y, 1) or (0, 1, 0)
 *                     like the reference's `cukernel_to_affine` output.
 */
#ifndef HALO2_ORACLE_H
#define HALO2_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* field selector: 0 = Fr (scalar field r), 1 = Fq (base field p) */
void or_field_mul(int fld, uint64_t* out, const uint64_t* a, const uint64_t* b, size_t n);
void or_field_add(int fld, uint64_t* out, const uint64_t* a, const uint64_t* b, size_t n);
void or_field_sub(int fld, uint64_t* out, const uint64_t* a, const uint64_t* b, size_t n);
void or_field_inv(int fld, uint64_t* out, const uint64_t* a, size_t n);
void or_field_to_mont(int fld, uint64_t* out, const uint64_t* a, size_t n);
void or_field_from_mont(int fld, uint64_t* out, const uint64_t* a, size_t n);
void or_fr_pow(uint64_t* out, const uint64_t* base, uint64_t e);
/* 64 little-endian bytes -> value mod r in Montgomery form (halo2curves from_u512) */
void or_fr_from_u512(uint64_t* out, const uint8_t* in, size_t n);

/* ChaCha20 keystream seeded like rand's `seed_from_u64`; element i takes block
 * (start + i) and is reduced with or_fr_from_u512. */
void or_chacha_seed_key(uint64_t seed, uint32_t key[8]);
void or_chacha_fr(uint64_t seed, uint64_t start, size_t n, uint64_t* out);
void or_chacha_block(const uint32_t key[8], uint64_t counter, uint8_t out[64]);

/* G1 */
void or_g1_generator(uint64_t* out_affine);
/* out[i] = s_i * G (affine); scalars in Montgomery form */
void or_g1_mul_gen(uint64_t* out_affine, const uint64_t* scalars, size_t n, int nthreads);
/* out = s * P, P affine; result in the 12-limb projective-out layout */
void or_g1_mul(uint64_t* out12, const uint64_t* p_affine, const uint64_t* s);
void or_g1_add_affine(uint64_t* out12, const uint64_t* p_affine, const uint64_t* q_affine);
int or_g1_is_on_curve(const uint64_t* p_affine);

/* Pippenger restatement of multiexp_serial + best_multiexp_cpu */
void or_msm(uint64_t* out12, const uint64_t* scalars, const uint64_t* points, size_t n, int nthreads);
/* double-and-add sum of s_i * P_i, independent of Pippenger (tiny n only) */
void or_msm_naive(uint64_t* out12, const uint64_t* scalars, const uint64_t* points, size_t n);

/* NTT family. All data in Montgomery form. */
void or_fft(uint64_t* a, uint32_t log_n, const uint64_t* omega, int nthreads);
void or_dft_naive(uint64_t* out, const uint64_t* a, uint32_t log_n, const uint64_t* omega);
void or_scale(uint64_t* a, size_t n, const uint64_t* c, int nthreads);
void or_distribute_powers(uint64_t* a, size_t n, const uint64_t* c, int nthreads);
void or_distribute_powers_zeta(uint64_t* a, size_t n, const uint64_t* coset_powers, uint32_t len);
void or_divide_by_vanishing(uint64_t* a, size_t n, const uint64_t* t_evals, uint32_t t_len);
void or_extended_from_lagrange_vec(uint64_t* out, const uint64_t* const* parts, uint32_t num_parts, size_t n);

/* products / scans / inversion */
void or_batch_invert(int fld, uint64_t* a, size_t n);
void or_grand_product(uint64_t* out, const uint64_t* in, const uint64_t* prefix, size_t n);
void or_lookup_product(uint64_t* out, const uint64_t* permuted_input, const uint64_t* permuted_table,
                       const uint64_t* compressed_input, const uint64_t* compressed_table,
                       const uint64_t* beta, const uint64_t* gamma, size_t n);
void or_permutation_product(uint64_t* modified, const uint64_t* const* perms, const uint64_t* const* values,
                            size_t batch, const uint64_t* beta, const uint64_t* gamma, const uint64_t* delta,
                            const uint64_t* omega, uint64_t* deltaomega, size_t n);
void or_quotient_lookups(uint64_t* values, const uint64_t* table_values, const uint64_t* product_coset,
                         const uint64_t* permuted_input_coset, const uint64_t* permuted_table_coset,
                         const uint64_t* l0, const uint64_t* l_last, const uint64_t* l_active_row,
                         const uint64_t* beta, const uint64_t* gamma, const uint64_t* y, size_t n);
void or_permutation_quotient(uint64_t* values, const uint64_t* const* perm_prod_cosets,
                             const uint64_t* const* perm_cosets, const uint64_t* const* column_values,
                             const uint64_t* l0, const uint64_t* l_last, const uint64_t* l_active_row,
                             const uint64_t* beta, const uint64_t* gamma, const uint64_t* y,
                             const uint64_t* delta, const uint64_t* delta_start,
                             const uint64_t* current_extended_omega, const uint64_t* omega,
                             size_t n_sets, size_t chunk_len, size_t n_perm_cols, int32_t last_rotation,
                             int32_t rot_scale, int32_t isize_, size_t n);

void or_quotient_rules(uint64_t* out, const uint64_t* prev, const uint64_t* const* fixed, size_t nf,
                       const uint64_t* const* instance, size_t ni, const uint64_t* const* advice, size_t na,
                       const uint64_t* challenges, const uint64_t* constants, const uint64_t* y,
                       const uint64_t* rules, size_t num_rules, const uint64_t* vp, size_t num_vp, size_t nq,
                       int64_t rot_scale, size_t row_start, size_t row_end);

void or_fr_inner_product(uint64_t* out, const uint64_t* a, const uint64_t* b, size_t n);

/* ---- opening-phase polynomial helpers ---- */
void or_eval_poly(uint64_t* out, const uint64_t* a, size_t n, const uint64_t* u);
void or_kate_division(uint64_t* q, const uint64_t* a, size_t n, const uint64_t* u);

/* ---- lookup permutation pair; returns 1 when an input value is missing from the table ---- */
int or_permute_expression_pair(uint64_t* pin, uint64_t* ptab, const uint64_t* input, const uint64_t* table,
                               size_t n, size_t U);

int or_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
