// bf16 vector types and the hi/lo split of an FP32 value shared by the BF16x3 kernels:
// x = hi + lo + O(2^-16 |x|) with hi = bf16_rne(x), lo = bf16_rne(x - hi).
#pragma once

#include <hip/hip_runtime.h>

namespace fc3 {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

static __device__ __forceinline__ void split_bf16(float x, __bf16& h, __bf16& l)
{
    h = static_cast<__bf16>(x);
    const float hf = static_cast<float>(h);
    l = static_cast<__bf16>(x - hf);
}

}  // namespace fc3
