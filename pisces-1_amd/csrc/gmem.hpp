// gmem.hpp -- global-address-space pointers for device code.
//
// Pointers that kernels read out of task descriptors are generic to the
// compiler, which then emits flat memory instructions: they count against both
// the vector-memory and the LDS counters, so LDS waits also wait for HBM loads
// and loads guarded by conditions become branches.  Casting such pointers to
// the global address space gives global_load / global_store instead.
#pragma once

#include <hip/hip_runtime.h>

namespace pisces {

template <typename T> using gptr = __attribute__((address_space(1))) T *;

template <typename T> __device__ __forceinline__ gptr<T> glb(T *p) { return (gptr<T>)p; }

// Constant address space: for data that no thread writes while the kernel
// runs (coefficients produced by earlier launches).  Wave-uniform loads
// through it become scalar loads; vector stores elsewhere in the kernel
// (possibly of the same type) no longer force them onto the vector path.
template <typename T> using cptr = __attribute__((address_space(4))) const T *;

template <typename T> __device__ __forceinline__ cptr<T> cst(const T *p) { return (cptr<T>)p; }

}  // namespace pisces
