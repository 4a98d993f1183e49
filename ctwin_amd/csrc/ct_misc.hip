// Small device kernels: register initialisation (engine::set, kernel.cl:514-520).
#include <hip/hip_runtime.h>

#include "ct_engine_impl.h"

namespace ct {

__global__ void k_set(int32_t *x, uint32_t n, uint32_t lanes, int32_t a)
{
	const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= (size_t)n * lanes) return;
	x[k] = ((k % n) == 0) ? a : 0;
}

hipError_t launch_set(int32_t *x, uint32_t n, uint32_t lanes, int32_t a, hipStream_t st)
{
	const size_t total = (size_t)n * lanes;
	const uint32_t blocks = (uint32_t)((total + 255) / 256);
	hipLaunchKernelGGL(k_set, dim3(blocks), dim3(256), 0, st, x, n, lanes, a);
	return hipGetLastError();
}

}   // namespace ct
