/*
 * halo2_gpu.h - C ABI of the MI355X (gfx950) hot path of the halo2-axiom-gpu
 * KZG prover.
 *
 * Every `_halo2_*` symbol below replaces the entry point of the same name in
 * the reference's `extern "C"` block (halo2_proofs/src/cuda/culib.rs, cited
 * per symbol), with the same argument order and meaning. A Rust caller keeps
 * its `#[link(name = "halo2_gpu")]` declarations; only `cudaStream_t` becomes a
 * HIP stream (`void*` here, `*mut c_void` on the Rust side).
 *
 * Conventions (as in the reference, culib.rs:1-10 and
 * cuda/include/common/halo2_ffi.h:8-60):
 *  - Field elements (Fr, Fq) are 32 bytes: 4 little-endian u64 limbs holding
 *    a * 2^256 mod m (Montgomery form), fully reduced.
 *  - G1 affine points are 64 bytes (x, y); the identity is (0, 0). MSM results
 *    are 96 bytes (x, y, 1) or the identity (0, 1, 0), in Montgomery form.
 *  - `void*` data arguments are caller-owned device pointers unless the name or
 *    the comment says host. `FFITraitObject*` points at one `{uint64_t* ptr}`
 *    word, or at an array of them for "many" entry points.
 *  - Launchers enqueue on `stream` and never call hipSetDevice. Entry points
 *    whose result lands in host memory synchronise `stream` before returning;
 *    device-output entry points do not.
 *  - Scratch memory is caller-owned: `_workspace_size` returns the bytes the
 *    matching launcher carves from `scratch`; a short workspace is reported as
 *    an error before anything is enqueued. The library never allocates device
 *    memory on these paths.
 *  - Errors come back as a 16-byte `RustError` by value: code 0 = success;
 *    otherwise `message` is a malloc'd string the caller frees with free().
 */
#ifndef HALO2_GPU_H
#define HALO2_GPU_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* cuda/include/common/halo2_ffi.h:30-48 ; src/cuda/error.rs:17-35 */
typedef struct RustError {
    int32_t code;
    char* message;
} RustError;

/* cuda/include/common/halo2_ffi.h:51-60 ; src/cuda/utils.rs:98-152 */
typedef struct FFITraitObject {
    uint64_t* ptr;
} FFITraitObject;

/* NTT kinds, cuda/src/ntt.cu:21-28 and src/poly/domain.rs:104-116 */
enum {
    HALO2_NTT_FFT = 1,
    HALO2_NTT_IFFT = 2,
    HALO2_NTT_COSET_FFT = 3,
    HALO2_NTT_ICOSET_FFT = 4,
    HALO2_NTT_IFFT_COSET_FFT = 5, /* unsupported, as in the reference */
    HALO2_NTT_COSET_FFT_PART = 6,
};

/* =========================================================== MSM (G1) */

/* culib.rs:30 ; cuda/src/multiexp.cu:166-171 */
size_t _halo2_msm_max_length(uint64_t free_bytes);
/* culib.rs:32 ; multiexp.cu:174-178 */
uint64_t _halo2_multiexp_workspace_size(uint64_t length);
/* culib.rs:35-43 ; multiexp.cu:258-335 — host scalars, host bases, host output */
RustError _halo2_multiexp(const FFITraitObject* scalar, const FFITraitObject* point,
    const FFITraitObject* output, size_t length, void* scratch, uint64_t scratch_bytes, void* stream);
/* culib.rs:45 ; multiexp.cu:341-345 */
uint64_t _halo2_multiexp_device_bases_workspace_size(uint64_t length);
/* culib.rs:48-56 ; multiexp.cu:355-424 — host scalars, device bases, host output */
RustError _halo2_multiexp_device_bases(const FFITraitObject* h_scalar, const void* d_bases,
    const FFITraitObject* h_output, size_t length, void* scratch, uint64_t scratch_bytes, void* stream);
/* culib.rs:58 ; multiexp.cu:431-435 */
uint64_t _halo2_multiexp_device_scalars_device_bases_workspace_size(uint64_t length);
/* culib.rs:62-70 ; multiexp.cu:447-525 — device scalars, device bases, host output */
RustError _halo2_multiexp_device_scalars_device_bases(const void* d_scalar, const void* d_bases,
    const FFITraitObject* h_output, size_t length, void* scratch, uint64_t scratch_bytes, void* stream);

/* No counterpart in the reference. For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): _halo2_multiexp_device_scalars_device_bases without the workspace
 * argument. Its scratch is a 64 MiB device array built into the library, one per
 * GPU, held until the process ends; lengths above h2amd_multiexp_noalloc_max_length()
 * (2^16 points) are rejected with an error. Calls that overlap on one GPU share that
 * array and race. Allocates nothing: no hipMalloc/hipFree. */
uint64_t h2amd_multiexp_noalloc_max_length(void);
RustError h2amd_multiexp_device_scalars_device_bases_noalloc(const void* d_scalar, const void* d_bases,
    const FFITraitObject* h_output, size_t length, void* stream);

/* ================================================================ NTT */

/* culib.rs:492 ; cuda/src/ntt.cu:383-393 */
uint64_t _halo2_fft_normal_workspace_size(uint32_t ntt_type, uint32_t log_n, uint32_t extend_log_n);
/* culib.rs:494-504 ; ntt.cu:395-474 — host input/output/omega/divisor; FFT, iFFT, cosetFFT,
 * icosetFFT and CosetFFT_Part at size 2^log_n */
RustError _halo2_fft_normal(uint32_t ntt_type, uint32_t log_n, const void* input, void* output,
    const void* omega, const void* divisor, void* scratch, uint64_t scratch_bytes, void* stream);
/* culib.rs:506-511 ; ntt.cu:149-164 — always true (deterministic sizing) */
bool _halo2_fft_normal_check_memory(uint32_t ntt_type, const void* input, uint32_t log_n,
    uint32_t extend_log_n);
/* culib.rs:513-517 ; ntt.cu:480-487 */
uint64_t _halo2_fft_normal_to_device_workspace_size(uint32_t ntt_type, uint32_t log_n,
    uint32_t extend_log_n);
/* culib.rs:522-532 ; ntt.cu:509-572 — device input/output (may alias), device omega,
 * host divisor; FFT, iFFT and CosetFFT_Part only */
RustError _halo2_fft_normal_to_device(uint32_t ntt_type, uint32_t log_n, const void* d_input,
    void* d_output, const void* d_omega, const void* h_divisor, void* scratch,
    uint64_t scratch_bytes, void* stream);
/* culib.rs:534-545 ; ntt.cu:576-639 — host input (2^log_n) -> host output (2^extend_log_n):
 * zero-pad, multiply index i by zeta^(i mod 3), DFT with the extended omega */
RustError _halo2_cosetfft(uint32_t ntt_type, uint32_t log_n, uint32_t extend_log_n, const void* input,
    void* output, const void* omega, const void* divisor, void* scratch, uint64_t scratch_bytes,
    void* stream);
/* culib.rs:547 ; ntt.cu:749-760 */
uint64_t _halo2_fft_many_workspace_size(uint32_t ntt_type, uint32_t log_n, uint32_t extend_log_n);
/* culib.rs:549-553 ; ntt.cu:765-775 */
uint64_t _halo2_fft_many_to_device_workspace_size(uint32_t ntt_type, uint32_t log_n,
    uint32_t extend_log_n);
/* culib.rs:560-572 ; ntt.cu:813-886 — host inputs/outputs, device omega, host divisor */
RustError _halo2_fft_many(uint32_t ntt_type, uint32_t num_many, uint32_t log_n, uint32_t extend_log_n,
    const FFITraitObject* input, const FFITraitObject* output, const void* d_omega,
    const FFITraitObject* divisor, void* scratch, uint64_t scratch_bytes, void* stream);
/* culib.rs:578-591 ; ntt.cu:898-962 — device outputs; inputs host or device per flag */
RustError _halo2_fft_many_to_device(uint32_t ntt_type, uint32_t num_many, uint32_t log_n,
    uint32_t extend_log_n, const FFITraitObject* input, const FFITraitObject* output,
    const void* d_omega, const FFITraitObject* divisor, bool input_on_device, void* scratch,
    uint64_t scratch_bytes, void* stream);
/* culib.rs:593 ; ntt.cu:780-786 — always true */
bool _halo2_fft_many_check_memory(uint32_t ntt_type, uint32_t log_n, uint32_t extend_log_n);
/* culib.rs:595-600 ; ntt.cu:666-681 — smallest split (log2) whose working set fits, -1 if none */
int32_t _halo2_get_fft_split_radix(uint32_t ntt_type, uint32_t log_n, uint32_t extend_log_n,
    uint64_t free_bytes);

/* ================================================ polynomial helpers */

/* culib.rs:383-389 ; cuda/src/polynomial/polynomial.cu:247-267
 * d_out[i * num_parts + p] = d_parts[p][i]; d_parts is a device array of device pointers */
RustError _halo2_extended_from_lagrange_vec_device(void* d_out, const void* d_parts, uint32_t num_parts,
    uint64_t n, void* stream);
/* culib.rs:311-317 ; polynomial.cu:270-290 — d_poly[i] *= d_t_evals[i % t_len] */
RustError _halo2_divide_by_vanishing_poly(void* d_poly, const void* d_t_evals, uint32_t t_len,
    uint64_t n, void* stream);
/* culib.rs:393-399 ; polynomial.cu:294-318
 * d_a[i] *= d_coset_powers[(i % (len + 1)) - 1] when i % (len + 1) != 0 */
RustError _halo2_distribute_powers_zeta(void* d_a, const void* d_coset_powers, uint32_t coset_powers_len,
    uint64_t n, void* stream);

/* ================================== products, scans and batch inversion */

/* culib.rs:72-80 ; cuda/src/polynomial/operations.cu:16-42 — in place x -> 1/x, 0 -> 0;
 * field_type 0 = Fr, 1 = Fq. Inputs are canonical (< m), as in the reference; during the
 * call the two top bits of the elements carry intermediate values (the result is canonical). */
RustError _halo2_batch_invert(void* d_data, uint32_t field_type, size_t length, void* stream);

/* culib.rs:127 ; cuda/src/commitment/permute_expression_pair.cu:202-227 */
uint64_t _halo2_permute_expression_pair_workspace_size(uint64_t n, uint64_t usable_rows);
/* culib.rs:129-141 ; permute_expression_pair.cu:229-350 — device in/out, n rows each: the
 * lookup permutation pair of src/plonk/lookup/prover.rs:788-882 (sorted input by canonical
 * value; run heads take their own value; leftover table values ascending fill the repeated
 * rows from the last one backwards; rows [usable_rows, n) zero). Error if usable_rows > n. */
RustError _halo2_permute_expression_pair(const void* d_compressed_input, const void* d_compressed_table,
    void* d_permuted_input, void* d_permuted_table, uint64_t n, uint64_t usable_rows, void* scratch,
    uint64_t scratch_bytes, void* stream);
/* culib.rs:81 ; cuda/src/commitment/commit_product.cu:33-43 */
size_t _halo2_commit_product_max_len(size_t length, uint64_t free_bytes);
/* culib.rs:82 ; commit_product.cu:45-48 */
uint64_t _halo2_commit_product_workspace_size(uint64_t poly_length);
/* culib.rs:87-100 ; commit_product.cu:50-130 — host polynomials at [offset, offset + length),
 * device beta/gamma: (A'+beta)(S'+gamma) inverted, times (A+beta)(S+gamma) */
RustError _halo2_commit_product(const FFITraitObject* lookup_product, const FFITraitObject* permuted_input,
    const FFITraitObject* permuted_table, const FFITraitObject* compressed_input,
    const FFITraitObject* compressed_table, const void* beta_device, const void* gamma_device,
    size_t length, size_t offset, void* scratch, uint64_t scratch_bytes, void* stream);
/* culib.rs:114-125 ; commit_product.cu:142-186 — every .ptr is a device pointer */
RustError _halo2_commit_product_device_inputs(const FFITraitObject* d_lookup_product,
    const FFITraitObject* d_permuted_input, const FFITraitObject* d_permuted_table,
    const FFITraitObject* d_compressed_input, const FFITraitObject* d_compressed_table,
    const void* beta_device, const void* gamma_device, size_t length, size_t offset, void* stream);

/* culib.rs:143 ; cuda/src/commitment/permutation_commit.cu:43-53 */
size_t _halo2_permutation_product_max_len(size_t length, uint64_t free_bytes);
/* culib.rs:144 ; permutation_commit.cu:55-58 */
uint64_t _halo2_permutation_product_workspace_size(uint64_t poly_length);
/* culib.rs:146-161 ; permutation_commit.cu:60-155 — host chunk buffers; modified_values is the
 * running accumulator (length entries); permutations/values are arrays of `batch` host columns
 * addressed at `offset`; deltaomega (device) is advanced by delta once per column */
RustError _halo2_permutation_product(const FFITraitObject* modified_values, const FFITraitObject* permutations,
    const FFITraitObject* values, const void* beta_device, const void* gamma_device, const void* delta_device,
    const void* omega_device, void* deltaomega_device, size_t length, size_t offset, size_t batch,
    void* scratch, uint64_t scratch_bytes, void* stream);
/* culib.rs:164 ; permutation_commit.cu:175-178 */
uint64_t _halo2_permutation_product_device_inputs_workspace_size(uint64_t poly_length);
/* culib.rs:180-195 ; permutation_commit.cu:180-256 — device variant */
RustError _halo2_permutation_product_device_inputs(const FFITraitObject* modified_values_device,
    const FFITraitObject* permutations_device, const FFITraitObject* values_device, const void* beta_device,
    const void* gamma_device, const void* delta_device, const void* omega_device, void* deltaomega_device,
    size_t length, size_t offset, size_t batch, void* scratch, uint64_t scratch_bytes, void* stream);

/* culib.rs:199 ; cuda/src/commitment/grand_product.cu:31-41 */
size_t _halo2_grand_product_max_len(size_t length, uint64_t free_bytes);
/* culib.rs:200 ; grand_product.cu:43-46 */
uint64_t _halo2_grand_product_workspace_size(uint64_t poly_length);
/* culib.rs:202-211 ; grand_product.cu:48-123 — output[i] = prefix * prod_{j<=i} input[j] over
 * [offset, offset + length), host buffers, host prefix */
RustError _halo2_grand_product(const FFITraitObject* output, const FFITraitObject* input,
    const FFITraitObject* prefix, size_t length, size_t offset, void* scratch, uint64_t scratch_bytes,
    void* stream);
/* culib.rs:221-227 ; grand_product.cu:125-195 — in place on device, device prefix */
RustError _halo2_grand_product_device_inputs(const FFITraitObject* d_inout, const FFITraitObject* d_prefix,
    size_t length, size_t offset, void* stream);

/* ===================================================== quotient pieces */

/* culib.rs:602-616 ; cuda/src/quotient_lookups.cu:16-53 — all device pointers */
RustError _halo2_quotient_lookups(void* values, const void* table_values, const void* product_coset,
    const void* permuted_input_coset, const void* permuted_table_coset, const void* l0, const void* l_last,
    const void* l_active_row, const void* beta, const void* gamma, const void* y, size_t poly_length,
    void* stream);
/* culib.rs:622-650 ; cuda/src/permutation_quotient.cu:23-93 — device vectors and pointer tables,
 * host scalars */
RustError _halo2_quotient_permutation(void* values, const void* l0, const void* l_last,
    const void* l_active_row, const void* perm_prod_cosets, const void* perm_cosets,
    const void* column_values, const void* beta, const void* gamma, const void* y, const void* delta,
    const void* delta_start, const void* current_extended_omega, const void* omega, uint64_t n_sets,
    uint64_t chunk_len, uint64_t n_perm_cols, int32_t last_rotation, int32_t rot_scale, int32_t isize_,
    uint64_t poly_length, void* stream);

/* Custom-gate quotient: compiled expression graph (128-bit CalcRules + 64-bit value-part
 * rules, src/plonk/evaluation.rs:107-350) interpreted per row over [row_start, row_end)
 * of an extended part of num_quotient rows; rotations wrap modulo num_quotient.
 * `rules`, `value_part_rules`, `challenges`, `constants` and `expr_constants`
 * ([0, 1, -1, 2, y]) are host memory. */
/* src/plonk/evaluation.rs:1435-1449 ; cuda/src/quotient.cu:309-331 */
uint64_t _halo2_evaluate_h_max_rows(const uint64_t* rules, uint64_t num_rules, uint64_t num_vp_rules,
    uint64_t num_quotient, uint64_t num_fixed, uint64_t num_instance, uint64_t num_advice,
    uint64_t num_challenges, uint64_t num_constants, uint64_t rotation_scale, uint64_t row_start,
    uint64_t row_end, uint64_t free_bytes);
/* evaluation.rs:1451-1464 ; quotient.cu:350-372 */
uint64_t _halo2_quotient_workspace_size(const uint64_t* rules, uint64_t num_rules, uint64_t num_vp_rules,
    uint64_t num_quotient, uint64_t num_fixed, uint64_t num_instance, uint64_t num_advice,
    uint64_t num_challenges, uint64_t num_constants, uint64_t rotation_scale, uint64_t row_start,
    uint64_t row_end);
/* evaluation.rs:1466-1490 ; quotient.cu:432-560 — host columns, host output */
RustError _halo2_quotient(const FFITraitObject* fixed, uint64_t num_fixed, const FFITraitObject* instance,
    uint64_t num_instance, const FFITraitObject* advices, uint64_t num_advice, FFITraitObject challenges,
    uint64_t num_challenges, FFITraitObject constants, uint64_t num_constants, FFITraitObject expr_constants,
    const uint64_t* rules, uint64_t num_rules, const uint64_t* value_part_rules, uint64_t num_vp_rules,
    FFITraitObject quotient_poly, uint64_t num_quotient, uint64_t rotation_scale, uint64_t row_start,
    uint64_t row_end, void* scratch, uint64_t scratch_bytes, void* stream);
/* evaluation.rs:1537-1560 — host arrays of device column pointers, host output */
RustError _halo2_quotient_device_columns(const void* const* fixed_d_ptrs, uint64_t num_fixed,
    const void* const* instance_d_ptrs, uint64_t num_instance, const void* const* advice_d_ptrs,
    uint64_t num_advice, FFITraitObject challenges, uint64_t num_challenges, FFITraitObject constants,
    uint64_t num_constants, FFITraitObject expr_constants, const uint64_t* rules, uint64_t num_rules,
    const uint64_t* value_part_rules, uint64_t num_vp_rules, FFITraitObject quotient_poly,
    uint64_t num_quotient, uint64_t rotation_scale, uint64_t row_start, uint64_t row_end, void* scratch,
    uint64_t scratch_bytes, void* stream);
/* evaluation.rs:1499-1523 — device columns, device output (row r - row_start), optional device
 * per-row Horner seed prev_values (null seeds zero) */
RustError _halo2_quotient_device_columns_device_out(const void* const* fixed_d_ptrs, uint64_t num_fixed,
    const void* const* instance_d_ptrs, uint64_t num_instance, const void* const* advice_d_ptrs,
    uint64_t num_advice, FFITraitObject challenges, uint64_t num_challenges, FFITraitObject constants,
    uint64_t num_constants, FFITraitObject expr_constants, const uint64_t* rules, uint64_t num_rules,
    const uint64_t* value_part_rules, uint64_t num_vp_rules, void* quotient_poly_device_ptr,
    uint64_t num_quotient, const void* prev_values_device_ptr, uint64_t rotation_scale, uint64_t row_start,
    uint64_t row_end, void* scratch, uint64_t scratch_bytes, void* stream);

/* ============================================= omega tables on device */

/* culib.rs:443 ; cuda/src/polynomial/omega.cu:72-76 */
uint64_t _halo2_generate_omega_powers_workspace_size(uint32_t log_n);
/* culib.rs:449-456 ; omega.cu:77-97 — d_omega_powers[i] = omega^i for i < 2^log_n */
RustError _halo2_generate_omega_powers(void* d_omega_powers, const void* d_omega, uint32_t log_n,
    void* scratch, uint64_t scratch_bytes, void* stream);

/* culib.rs:430 ; omega.cu:15-20 */
uint64_t _halo2_power_of_omega_workspace_size(uint32_t log_n);
/* culib.rs:432-441 ; omega.cu:29-70 — host omega in, host outputs: omega_lut[0] = 1,
 * omega_lut[i] = omega^(2^(i-1)) for i = 1..log_n, res = omega^pow. Error when log_n >= 32
 * or pow >= 2^log_n. Synchronises `stream`. */
RustError _halo2_power_of_omega(const FFITraitObject* res, const FFITraitObject* omega_lut,
    const FFITraitObject* omega, uint32_t log_n, uint32_t pow, void* scratch, uint64_t scratch_bytes,
    void* stream);
/* culib.rs:458 ; omega.cu:99-105 */
uint64_t _halo2_generate_omega_lut_workspace_size(uint32_t log_n);
/* culib.rs:460-467 ; omega.cu:107-151 — host output of 2^10 + 2^(log_n-10) entries:
 * omega^i for i < 2^10, then omega^(i * 2^10) for i < 2^(log_n - 10). log_n in [10, 42). */
RustError _halo2_generate_omega_lut(const FFITraitObject* sparse_twiddle, const FFITraitObject* omega,
    uint32_t log_n, void* scratch, uint64_t scratch_bytes, void* stream);
/* culib.rs:469-475 ; omega.cu:153-161 */
uint64_t _halo2_generate_omegadelta_workspace_size(uint32_t log_n, uint32_t omega_start, uint32_t omega_end,
    uint32_t colunm_num, uint32_t colunm_offset);
/* culib.rs:477-490 ; omega.cu:163-210, kernel/omega.h:226-260 — device output
 * out[i - omega_start] = omega^row * delta^col for i in [omega_start, omega_end], with
 * (col, row) = (mapping[2i], mapping[2i + 1]) a device u32 array; omega and delta are host
 * 32-byte scalars; delta powers are tabled up to the next power of two >= colunm_num.
 * log_n in [10, 42). colunm_offset is accepted and unused, as in the reference. */
RustError _halo2_generate_omegadelta(void* omegadelta, const void* mapping, const void* omega_host,
    const void* delta_host, uint32_t log_n, uint32_t omega_start, uint32_t omega_end, uint32_t colunm_num,
    uint32_t colunm_offset, void* scratch, uint64_t scratch_bytes, void* stream);

/* ================================= opening-phase polynomial helpers */

/* culib.rs:231 ; cuda/src/polynomial/polynomial.cu:76-79 */
uint64_t _halo2_eval_polynomial_workspace_size(uint64_t length);
/* culib.rs:233-243 ; polynomial.cu:81-108 — p(point) for a device poly of `length`
 * coefficients and a host point; the 32-byte result stays in `scratch` and its device
 * address is written to *d_result_out. */
RustError _halo2_eval_polynomial(const void* d_poly, const FFITraitObject* point, void** d_result_out,
    uint64_t length, void* scratch, uint64_t scratch_bytes, void* stream);
/* culib.rs:245-249 ; polynomial.cu:140-152 — largest power-of-two halving of poly_length whose
 * batch workspace fits free_bytes (0 if none) */
size_t _halo2_eval_poly_batch_max_len(size_t poly_length, size_t batch_size, uint64_t free_bytes);
/* culib.rs:251 ; polynomial.cu:154-160 */
uint64_t _halo2_eval_polynomial_batch_workspace_size(uint64_t poly_length, uint64_t batch_size);
/* culib.rs:253-264 ; polynomial.cu:162-245 — host polys (coefficients
 * [poly_offset, poly_offset + poly_length) of each), host points, host results:
 * eval_result[i] = sum_j poly_i[poly_offset + j] * point_i^j. Synchronises `stream`. */
RustError _halo2_eval_polynomial_batch(const FFITraitObject* poly_in_many, const FFITraitObject* eval_points,
    const FFITraitObject* eval_result, size_t poly_offset, size_t poly_length, size_t batch_size, void* scratch,
    uint64_t scratch_bytes, void* stream);
/* No counterpart in the reference (it evaluates one device polynomial per
 * _halo2_eval_polynomial call, src/cuda/polynomial_ops.rs:142-222): count device
 * polynomials (d_polys[i] with lengths[i] coefficients) at host points (32-byte
 * Montgomery elements) into d_results[i] (device, 32 bytes each), in one pair of launches
 * per 32 polynomials; stream-ordered, no synchronisation. */
uint64_t h2amd_eval_polynomial_batch_device_workspace_size(const uint64_t* lengths, uint32_t count);
RustError h2amd_eval_polynomial_batch_device(const void* const* d_polys, const uint64_t* lengths,
    const void* h_points, uint32_t count, void* d_results, void* scratch, uint64_t scratch_bytes, void* stream);
/* culib.rs:401-405 ; cuda/src/multiopen.cu:47-57 */
size_t _halo2_multiopen_poly_max_len(size_t poly_length, size_t batch_size, uint64_t free_bytes);
/* culib.rs:407-410 ; multiopen.cu:65-71 */
uint64_t _halo2_multiopen_poly_calculation_workspace_size(uint64_t poly_length, uint64_t batch_size);
/* culib.rs:412-426 ; multiopen.cu:76-171 — host inputs/outputs over coefficients
 * [poly_offset, poly_offset + poly_length): poly_acc = sum_i challenge_i * poly_i (the slice
 * of poly_acc is overwritten) and evaluate_result[i] = slice of poly_i at evaluate_point[i].
 * Synchronises `stream`. */
RustError _halo2_multiopen_poly_calculation(const FFITraitObject* poly_in_many, const FFITraitObject* poly_acc,
    size_t poly_offset, size_t poly_length, size_t batch_size, const FFITraitObject* challenge_point,
    const FFITraitObject* evaluate_point, const FFITraitObject* evalaute_result, void* scratch,
    uint64_t scratch_bytes, void* stream);
/* culib.rs:266-274 ; polynomial.cu:435-455 — d_poly_acc[i] += *d_scalar * d_poly_in[i] */
RustError _halo2_poly_multiply_add(void* d_poly_acc, const void* d_poly_in, const void* d_scalar,
    size_t poly_length, void* stream);
/* No counterpart in the reference (it folds one device polynomial per
 * _halo2_poly_multiply_add call, src/cuda/polynomial_ops.rs:397-439, as the SHPLONK prover
 * does for every rotation set, shplonk/prover.rs:183-268): d_acc[i] += sum_j
 * d_scalars[j] * d_polys[j][i] for i < n, where d_polys[j] (host array of device pointers)
 * holds lengths[j] <= n coefficients and is zero beyond them; d_scalars is a device table
 * of count 32-byte Montgomery elements. One pass over d_acc per 16 polynomials;
 * stream-ordered, no synchronisation. */
RustError h2amd_poly_multiply_add_many(void* d_acc, const void* const* d_polys, const uint64_t* lengths,
    const void* d_scalars, uint32_t count, uint64_t n, void* stream);
/* culib.rs:276-283 ; polynomial.cu:321-341 — d_out[i] = d_a[i] * d_b[i] (aliasing allowed) */
RustError _halo2_poly_elementwise_multiply(void* d_out, const void* d_a, const void* d_b, uint64_t length,
    void* stream);
/* culib.rs:285-290 ; polynomial.cu:343-361 — d_out[i] = 1 (Montgomery one) */
RustError _halo2_poly_fill_one(void* d_out, uint64_t length, void* stream);
/* culib.rs:352-358 ; polynomial.cu:363-385 — d_acc[i] -= d_short[i], i < short_len */
RustError _halo2_poly_sub_short_inplace(void* d_acc, const void* d_short, uint64_t short_len, void* stream);
/* culib.rs:360-371 ; polynomial.cu:390-416 — d_out[i] = d_long[i] - d_short[i] for
 * i < short_len, d_long[i] for short_len <= i < long_len */
RustError _halo2_poly_sub_short_out_of_place(void* d_out, const void* d_long, const void* d_short,
    uint64_t short_len, uint64_t long_len, void* stream);
/* culib.rs:373-380 ; polynomial.cu:418-433 — d_buf[0] -= *d_scalar */
RustError _halo2_poly_sub_scalar_at_zero(void* d_buf, const void* d_scalar, void* stream);

/* culib.rs:319 ; cuda/src/polynomial/kate_division.cu:27-36 */
uint64_t _halo2_kate_division_workspace_size(uint64_t n);
/* culib.rs:321-333 ; kate_division.cu:49-135 — d_q (n - 1 coefficients) =
 * (a(X) - a(u)) / (X - u) for the n coefficients at d_a and the device scalar d_u;
 * n < 2 is a no-op. */
RustError _halo2_kate_division_device(const void* d_a, void* d_q, const void* d_u, uint64_t n, void* scratch,
    uint64_t scratch_bytes, void* stream);
/* culib.rs:335-347 ; kate_division.cu:148-250 — same quotient in d_q[0, n - 1), zeros in
 * [n - 1, out_len); with out_len < n - 1 only the first out_len coefficients are written. */
RustError _halo2_kate_division_device_padded(const void* d_a, void* d_q, const void* d_u, uint64_t n,
    uint64_t out_len, void* scratch, uint64_t scratch_bytes, void* stream);

/* culib.rs:20-26 ; cuda/include/kernel/decode_assigned.h:10-14 — byte layout of one
 * Assigned<F> element: tag byte at offset 0 (0 Zero, 1 Trivial, 2 Rational), numerator
 * and denominator payloads at the given offsets */
typedef struct AssignedLayout {
    uint32_t stride_bytes;
    uint32_t num_offset;
    uint32_t denom_offset;
} AssignedLayout;
/* culib.rs:292-308 ; cuda/src/polynomial/decode_assigned.cu:28-53 — d_nums[i], d_denoms[i] =
 * (0, 1) / (x, 1) / (n, d) for Zero / Trivial(x) / Rational(n, d) at d_raw + i * stride.
 * Payload offsets must be 4-byte aligned and leave 32 bytes inside the stride. */
RustError _halo2_decode_assigned(void* d_nums, void* d_denoms, const void* d_raw, uint64_t n,
    AssignedLayout layout, void* stream);

/* ======================== device data generation (no reference counterpart)
 * Helpers used by the benchmark and tests to build large seeded inputs in HBM
 * without a host round trip. */

/* out[i] = Fr from ChaCha20 block (start + i) of the stream seeded like rand's
 * `seed_from_u64(seed)` (64 uniform bytes reduced mod r, Montgomery form). */
RustError h2amd_random_fr(void* d_out, uint64_t n, uint64_t seed, uint64_t start, void* stream);
/* d_out[i] = d_scalars[i] * G as 64-byte affine points (identity -> (0, 0)); scalars in
 * Montgomery form, G = (1, 2). Builds KZG SRS points tau^i * G and s_i * G test bases. */
uint64_t h2amd_g1_mul_generator_workspace_size(uint64_t n);
RustError h2amd_g1_mul_generator(void* d_out, const void* d_scalars, uint64_t n, void* scratch,
    uint64_t scratch_bytes, void* stream);

/* ================== fixed-base MSM (no reference counterpart)
 * For bases that stay resident (the KZG SRS): a table of rows 2^(c w) P_i (affine,
 * internal 2^261 scale, 64 B per point, W rows of n points) lets every window share
 * one bucket set, removing the per-window bucket reduction and the Horner over
 * windows. Results equal `_halo2_multiexp_device_scalars_device_bases` bit for bit. */

/* bytes of the table for n bases (W(n) rows of n points) */
uint64_t h2amd_msm_table_bytes(uint64_t n);
uint64_t h2amd_msm_precompute_workspace_size(uint64_t n);
/* d_table (h2amd_msm_table_bytes(n) bytes) from n device affine bases */
RustError h2amd_msm_precompute(void* d_table, const void* d_bases, uint64_t n, void* scratch,
    uint64_t scratch_bytes, void* stream);
uint64_t h2amd_multiexp_precomputed_workspace_size(uint64_t length, uint64_t table_n);
/* sum_i s_i P_i over the first `length` <= table_n points of a table built for table_n
 * points; device scalars, host 96-byte output (as `_halo2_multiexp_*`). Synchronises. */
RustError h2amd_multiexp_precomputed(const void* d_scalar, const void* d_table, uint64_t table_n,
    const FFITraitObject* h_output, uint64_t length, void* scratch, uint64_t scratch_bytes, void* stream);

/* ======================= uneven NTT batches (no reference counterpart)
 * `count` transforms of one kind (FFT=1, iFFT=2, cosetFFT=3, CosetFFT_Part=6; as
 * `_halo2_fft_many_to_device` with device input) but each with its own log_n /
 * extend_log_n, omega (device word) and divisor (host word): every pass of all of them
 * runs in one launch, so small transforms fill the chip beside large ones (the
 * coeff_to_extended of all polynomials of a proof, which the reference runs per size).
 * Outputs must not alias inputs; results equal the per-transform entry points. */
uint64_t h2amd_fft_batch_to_device_workspace_size(uint32_t ntt_type, uint32_t count, const uint32_t* log_n,
    const uint32_t* extend_log_n);
RustError h2amd_fft_batch_to_device(uint32_t ntt_type, uint32_t count, const uint32_t* log_n,
    const uint32_t* extend_log_n, const void* const* d_in, void* const* d_out, const void* const* d_omega,
    const FFITraitObject* divisors, void* scratch, uint64_t scratch_bytes, void* stream);

/* ======================= batched MSMs (no reference counterpart)
 * Up to 64 MSMs of uneven lengths over one shared device base array, in one sort /
 * accumulate / reduce pipeline: MSM j = sum_{i < lengths[j]} s_j[i] * bases[i] (the
 * KZG commitments of many polynomials against one SRS, which the reference runs as
 * one `_halo2_multiexp_device_scalars_device_bases` call each, multiexp.rs:215-256).
 * d_scalars: `count` device pointers (host array); d_bases: at least max(lengths)
 * affine bases; h_out: count x 96 bytes, each equal bit for bit to the single-MSM
 * result. Synchronises. */
uint64_t h2amd_multiexp_batch_workspace_size(const uint64_t* lengths, uint32_t count);
RustError h2amd_multiexp_batch(const void* const* d_scalars, const uint64_t* lengths, uint32_t count,
    const void* d_bases, void* h_out, void* scratch, uint64_t scratch_bytes, void* stream);
/* The same batch over a precomputed table (h2amd_msm_precompute of table_n bases, rows
 * 2^(c w) P): every lengths[j] <= table_n; each MSM's windows share one bucket set, so
 * there is no doubling chain per MSM. Results equal the single-MSM ones bit for bit.
 * Synchronises. */
uint64_t h2amd_multiexp_batch_precomputed_workspace_size(const uint64_t* lengths, uint32_t count, uint64_t table_n);
RustError h2amd_multiexp_batch_precomputed(const void* const* d_scalars, const uint64_t* lengths, uint32_t count,
    const void* d_table, uint64_t table_n, void* h_out, void* scratch, uint64_t scratch_bytes, void* stream);

/* ======================= SRS point formats (G1 halves of ParamsKZG::write_custom /
 * read_custom, src/poly/kzg/commitment.rs:317-414; no FFI counterpart: the reference
 * decodes on the host). Points are device affine 64-byte (x, y) Montgomery words.
 * Validation (SerdeFormat::RawBytes): a coordinate >= p or a point off y^2 = x^3 + 3 is
 * invalid; the identity (0, 0) is valid. Compressed (SerdeFormat::Processed): 32 bytes,
 * canonical x little-endian with the parity of canonical y in bit 7 of byte 31, all-zero
 * = identity. Validation and decompression report the invalid count and the first
 * invalid index (~0 when none) in host words and synchronise; they need
 * h2amd_g1_flags_workspace_size() bytes of device workspace. */
uint64_t h2amd_g1_flags_workspace_size(void);
RustError h2amd_g1_validate(const void* d_points, uint64_t n, uint64_t* h_invalid, uint64_t* h_first_invalid,
    void* scratch, uint64_t scratch_bytes, void* stream);
RustError h2amd_g1_compress(void* d_out, const void* d_points, uint64_t n, void* stream);
RustError h2amd_g1_decompress(void* d_points, const void* d_in, uint64_t n, uint64_t* h_invalid,
    uint64_t* h_first_invalid, void* scratch, uint64_t scratch_bytes, void* stream);

/* ======================= Lagrange bases from coefficient bases (no FFI counterpart:
 * the reference runs halo2-axiom's g_to_lagrange on the host, from ParamsKZG::from_parts and
 * downsize, src/poly/kzg/commitment.rs:205-228 and :436-450). d_out[i] = n^-1 sum_j
 * omega_inv^(i j) d_g[j] for n = 2^k device affine points, omega_inv (host Montgomery word)
 * the inverse of the 2^k-th root of unity; affine output, identity (0, 0). Enqueues only. */
uint64_t h2amd_g1_to_lagrange_workspace_size(uint32_t k);
RustError h2amd_g1_to_lagrange(void* d_out, const void* d_g, uint32_t k, const void* h_omega_inv, void* scratch,
    uint64_t scratch_bytes, void* stream);

/* ======================= tuned MSM window widths (no reference counterpart)
 * widths[i] (0 = the default rule) for i < count <= 64 replaces the whole table: entry
 * i < 32 (4..20) is the window width of plain MSMs of 2^i .. 2^(i+1) - 1 points, entry
 * 32 + i (4..22) that of precomputed tables built for 2^i .. 2^(i+1) - 1 bases. It changes
 * workspace and table sizes too, so set it before sizing; a table is only valid under
 * the width it was built with (h2amd_msm_table_width(table_n) reports the current one).
 * Process-global, not synchronised with concurrent MSM calls. h2amd_msm_width(n) reports
 * the width a plain MSM of n points uses. The tuner is halo2_amd/autotune.py. */
RustError h2amd_msm_set_widths(const uint8_t* widths, uint32_t count);
uint32_t h2amd_msm_width(uint64_t n);
uint32_t h2amd_msm_table_width(uint64_t table_n);
/* tiles[log_n] (0 = default, else 6..12): log2 of the elements one NTT work-group
 * transforms (radix columns x column group) for transforms of 2^log_n; same rules. */
RustError h2amd_ntt_set_tiles(const uint8_t* tiles, uint32_t count);

/* ======================= kernel timing (no reference counterpart)
 * When enabled, launchers bracket their hot kernels ("ntt_pass", "msm_bucket_level0",
 * ...) with HIP events on the launch stream; reading syncs the recorded events. */
void h2amd_profile_enable(int on);
void h2amd_profile_reset(void);
int h2amd_profile_read(const char* kernel, uint64_t* launches, double* total_ms);
/* Measured v_mad_u64_u32 throughput (the integer multiply-add the field arithmetic
 * runs on); scratch holds blocks * 256 * 8 bytes. */
RustError h2amd_mad_peak_probe(void* scratch, uint32_t blocks, uint32_t iters, double* mads_per_s, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* HALO2_GPU_H */
