// SRS point formats on the device: the G1 halves of ParamsKZG::write_custom /
// read_custom (src/poly/kzg/commitment.rs:317-414 of the reference).
//
//   RawBytes / RawBytesUnchecked: a G1 point is its 64-byte in-memory form
//     (x, y as 4 little-endian 64-bit Montgomery limbs; identity (0, 0)), which
//     is also the device layout of the bases, so reading is one copy. The
//     checked form rejects a coordinate >= p or a point off y^2 = x^3 + 3:
//     g1_check_kernel counts such points and finds the first.
//   Processed: 32 bytes, x in canonical little-endian form with the parity of
//     canonical y in bit 7 of byte 31; all-zero bytes are the identity.
//     Decompression takes y = (x^3 + 3)^((p+1)/4) (p = 3 mod 4) and negates it
//     when its parity disagrees with the flag.
#include "common.h"
#include "curve.h"

using namespace h2;

namespace {

// (p + 1) / 4, little-endian 32-bit words
__constant__ uint32_t kSqrtExp[8] = { 0xb61f3f52u, 0x4f082305u, 0x5a1c72a3u, 0x65e05aa4u, 0xa0605617u, 0x6e14116du,
    0xb84c680au, 0x0c19139cu };

// a^e for a 256-bit exponent, 4-bit fixed windows
__device__ Fq pow_words(const Fq& a, const uint32_t* e)
{
    Fq tbl[16];
    tbl[0] = Fq::one();
    tbl[1] = a;
#pragma unroll
    for (int i = 2; i < 16; ++i)
        tbl[i] = Fq::mul(tbl[i - 1], a);
    Fq r = Fq::one();
    bool started = false;
    for (int w = 63; w >= 0; --w) {
        const uint32_t nib = (e[w >> 3] >> ((w & 7) * 4)) & 0xfu;
        if (started) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                r = Fq::sqr(r);
        }
        if (nib) {
            Fq s = tbl[0];
#pragma unroll
            for (int i = 1; i < 16; ++i)
                if ((uint32_t)i == nib)
                    s = tbl[i];
            r = started ? Fq::mul(r, s) : s;
            started = true;
        }
    }
    return r;
}

// x < p as 256-bit integers
__device__ __forceinline__ bool below_p(const Fq& x)
{
    for (int i = 7; i >= 0; --i) {
        if (x.v[i] != FqParams::M[i])
            return x.v[i] < FqParams::M[i];
    }
    return false;
}

__device__ __forceinline__ Fq curve_b() // 3 in Montgomery form
{
    const Fq one = Fq::one();
    return Fq::add(Fq::add(one, one), one);
}

__device__ __forceinline__ bool on_curve(const Fq& x, const Fq& y)
{
    const Fq rhs = Fq::add(Fq::mul(Fq::sqr(x), x), curve_b());
    return Fq::sqr(y) == rhs;
}

// flags[0] = invalid points, flags[1] = first invalid index (0xffffffff if none)
__global__ void flags_init_kernel(uint32_t* flags)
{
    if (threadIdx.x == 0) {
        flags[0] = 0;
        flags[1] = 0xffffffffu;
    }
}

__device__ __forceinline__ void flag_bad(uint32_t* flags, uint64_t i)
{
    atomicAdd(&flags[0], 1u);
    atomicMin(&flags[1], (uint32_t)i);
}

__global__ void g1_check_kernel(const uint8_t* __restrict__ pts, uint64_t n, uint32_t* __restrict__ flags)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const G1Affine p = G1Affine::load(pts + 64 * i);
    if (p.is_inf())
        return;
    if (!below_p(p.x) || !below_p(p.y) || !on_curve(p.x, p.y))
        flag_bad(flags, i);
}

__global__ void g1_compress_kernel(const uint8_t* __restrict__ pts, uint8_t* __restrict__ out, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const G1Affine p = G1Affine::load(pts + 64 * i);
    Fq x = Fq::zero();
    if (!p.is_inf()) {
        x = Fq::from_mont(p.x);
        const Fq y = Fq::from_mont(p.y);
        x.v[7] |= (y.v[0] & 1u) << 31;
    }
    x.store(out + 32 * i);
}

__global__ void g1_decompress_kernel(const uint8_t* __restrict__ in, uint8_t* __restrict__ pts, uint64_t n,
    uint32_t* __restrict__ flags)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    Fq x = Fq::load(in + 32 * i);
    const uint32_t ysign = x.v[7] >> 31;
    x.v[7] &= 0x7fffffffu;
    G1Affine p { Fq::zero(), Fq::zero() };
    if (!below_p(x)) {
        flag_bad(flags, i);
    } else if (!(x.is_zero() && !ysign)) {
        const Fq xm = Fq::to_mont(x);
        const Fq rhs = Fq::add(Fq::mul(Fq::sqr(xm), xm), curve_b());
        Fq y = pow_words(rhs, kSqrtExp);
        if (Fq::sqr(y) != rhs) {
            flag_bad(flags, i); // x^3 + 3 is not a square: no point has this x
        } else {
            if ((Fq::from_mont(y).v[0] & 1u) != ysign)
                y = Fq::neg(y);
            p = G1Affine { xm, y };
        }
    }
    p.store(pts + 64 * i);
}

unsigned blocks_for(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

RustError read_flags(uint32_t* d_flags, uint64_t* h_invalid, uint64_t* h_first, hipStream_t stream)
{
    uint32_t f[2];
    H2_TRY(hipMemcpyAsync(f, d_flags, 8, hipMemcpyDeviceToHost, stream));
    H2_TRY(hipStreamSynchronize(stream));
    if (h_invalid)
        *h_invalid = f[0];
    if (h_first)
        *h_first = f[0] ? (uint64_t)f[1] : ~0ull;
    return ok();
}

} // namespace

extern "C" {

uint64_t h2amd_g1_flags_workspace_size(void) { return 256; }

RustError h2amd_g1_validate(const void* d_points, uint64_t n, uint64_t* h_invalid, uint64_t* h_first_invalid,
    void* scratch, uint64_t scratch_bytes, void* stream_)
{
    hipStream_t stream = (hipStream_t)stream_;
    H2_CHECK_ARG(n < (1ull << 32), "h2amd_g1_validate: %lu points exceed 2^32", (unsigned long)n);
    H2_CHECK_ARG(scratch && scratch_bytes >= 8, "h2amd_g1_validate: needs 8 bytes of workspace");
    uint32_t* flags = (uint32_t*)scratch;
    flags_init_kernel<<<1, 64, 0, stream>>>(flags);
    H2_LAUNCH_CHECK();
    if (n) {
        g1_check_kernel<<<blocks_for(n, 256), 256, 0, stream>>>((const uint8_t*)d_points, n, flags);
        H2_LAUNCH_CHECK();
    }
    return read_flags(flags, h_invalid, h_first_invalid, stream);
}

RustError h2amd_g1_compress(void* d_out, const void* d_points, uint64_t n, void* stream_)
{
    hipStream_t stream = (hipStream_t)stream_;
    if (n == 0)
        return ok();
    g1_compress_kernel<<<blocks_for(n, 256), 256, 0, stream>>>((const uint8_t*)d_points, (uint8_t*)d_out, n);
    H2_LAUNCH_CHECK();
    return ok();
}

RustError h2amd_g1_decompress(void* d_points, const void* d_in, uint64_t n, uint64_t* h_invalid,
    uint64_t* h_first_invalid, void* scratch, uint64_t scratch_bytes, void* stream_)
{
    hipStream_t stream = (hipStream_t)stream_;
    H2_CHECK_ARG(n < (1ull << 32), "h2amd_g1_decompress: %lu points exceed 2^32", (unsigned long)n);
    H2_CHECK_ARG(scratch && scratch_bytes >= 8, "h2amd_g1_decompress: needs 8 bytes of workspace");
    uint32_t* flags = (uint32_t*)scratch;
    flags_init_kernel<<<1, 64, 0, stream>>>(flags);
    H2_LAUNCH_CHECK();
    if (n) {
        g1_decompress_kernel<<<blocks_for(n, 256), 256, 0, stream>>>((const uint8_t*)d_in, (uint8_t*)d_points, n,
            flags);
        H2_LAUNCH_CHECK();
    }
    return read_flags(flags, h_invalid, h_first_invalid, stream);
}

} // extern "C"
