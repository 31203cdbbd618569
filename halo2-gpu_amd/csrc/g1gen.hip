// Fixed-base G1 scalar multiplication s_i * G on the device, used to build
// seeded MSM inputs and KZG SRS points (G_i = tau^i * G) directly in HBM.
// A table of 2^j * G (j < 254, affine) is built once per call; each thread then
// adds the table entries selected by the bits of its canonical scalar (XYZZ
// mixed additions) and converts the sum to affine.
#include "common.h"
#include "curve.h"

using namespace h2;

namespace {

// table[j] = 2^j * G for j < 254, G = (1, 2)
__global__ void gen_table_xyzz_kernel(G1Xyzz* __restrict__ tab)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    Fq one = Fq::one();
    G1Affine g { one, Fq::dbl(one) };
    G1Xyzz cur = G1Xyzz::from_affine(g);
    for (int j = 0; j < 254; ++j) {
        cur.store(tab + j);
        cur = G1Xyzz::dbl(cur);
    }
}

__global__ void table_to_affine_kernel(const G1Xyzz* __restrict__ tab, uint8_t* __restrict__ aff)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < 254)
        G1Xyzz::load(tab + j).to_affine().store(aff + 64 * j);
}

__global__ void __launch_bounds__(256) mul_generator_kernel(const Fr* __restrict__ scal,
    const uint8_t* __restrict__ tab, uint8_t* __restrict__ out, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const Fr s = Fr::from_mont(Fr::load(scal + i));
    G1Xyzz acc = G1Xyzz::inf();
    for (int j = 0; j < 254; ++j)
        if ((s.v[j >> 5] >> (j & 31)) & 1u)
            acc = G1Xyzz::add_affine(acc, G1Affine::load(tab + 64 * j));
    acc.to_affine().store(out + 64 * i);
}

} // namespace

extern "C" uint64_t h2amd_g1_mul_generator_workspace_size(uint64_t n)
{
    (void)n;
    return 254 * 128 + 254 * 64 + 1024;
}

extern "C" RustError h2amd_g1_mul_generator(void* d_out, const void* d_scalars, uint64_t n, void* scratch,
    uint64_t scratch_bytes, void* stream_)
{
    hipStream_t stream = (hipStream_t)stream_;
    if (n == 0)
        return ok();
    Workspace ws(scratch, scratch_bytes);
    G1Xyzz* tx = (G1Xyzz*)ws.take(254 * 128);
    uint8_t* ta = (uint8_t*)ws.take(254 * 64);
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "h2amd_g1_mul_generator: workspace of %lu bytes is too small",
            (unsigned long)scratch_bytes);
    gen_table_xyzz_kernel<<<1, 64, 0, stream>>>(tx);
    H2_LAUNCH_CHECK();
    table_to_affine_kernel<<<4, 64, 0, stream>>>(tx, ta);
    H2_LAUNCH_CHECK();
    mul_generator_kernel<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>((const Fr*)d_scalars, ta,
        (uint8_t*)d_out, n);
    H2_LAUNCH_CHECK();
    return ok();
}
