// GPU window sieve (placeholder until the kernels land).
#include <hip/hip_runtime.h>

#include "../../include/ctwin.h"
#include "ct_engine_impl.h"

int ct_sieve_window(int, int, uint64_t, uint64_t, uint64_t, uint64_t, uint64_t, uint64_t, int, uint8_t *)
{
	return ct::fail(CT_ERR_UNSUPPORTED, "sieve not built yet");
}
