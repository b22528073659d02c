// Matrix-core ceiling of the device, measured with the instruction the
// learned corrector's convolutions issue (v_mfma_f32_16x16x32_bf16).  This
// file is compiled on its own with -mllvm -amdgpu-mfma-vgpr-form: without
// it the compiler keeps the eight accumulators in AGPRs and rotates them
// through VGPRs every iteration (56 moves per 8 MFMA), which would time the
// moves rather than the matrix cores.  The convolution kernels already use
// the VGPR form.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nmr_kernels.h"

namespace nmr {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float fl4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x8 as_bf8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// Every wave runs `iters` rounds of 8 independent accumulation chains
// (16 384 FLOP per instruction) on register operands, with no memory
// traffic in the loop; one float per thread keeps the accumulators live.
__global__ void __launch_bounds__(256) mfma_rate_kernel(int iters, float* __restrict__ out)
{
    const uint32_t t = threadIdx.x + 1u;
    const bf16x8 A = as_bf8(make_uint4(0x3f803f80u ^ t, 0x3f803f80u, 0x3f803f80u, 0x3f803f80u ^ (t << 4)));
    const bf16x8 B = as_bf8(make_uint4(0x3c003c00u, 0x3c003c00u ^ t, 0x3c003c00u, 0x3c003c00u));
    fl4 acc[8];
#pragma unroll
    for (int c = 0; c < 8; c++) acc[c] = fl4{0.f, 0.f, 0.f, (float)c};
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int c = 0; c < 8; c++) acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[c], 0, 0, 0);
    }
    float sum = 0.f;
#pragma unroll
    for (int c = 0; c < 8; c++) sum += (acc[c][0] + acc[c][1]) + (acc[c][2] + acc[c][3]);
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = sum;
}

} // namespace

hipError_t launch_mfma_rate(int iters, int grid, float* out, hipStream_t s)
{
    hipLaunchKernelGGL(mfma_rate_kernel, dim3(grid), dim3(256), 0, s, iters, out);
    return hipGetLastError();
}

} // namespace nmr
