// gmem.hpp -- address-space pointers and FP64 reciprocals for device code.
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

// Loads of block columns that one workgroup of a launch streams through once
// carry the nontemporal hint (the nt bit): on MI355X a plain read stream of
// 1 GiB measured 7.1 TB/s with it against 6.3 TB/s with the default policy
// (pisces_stream_probe).  Columns that several workgroups of the same launch
// read again from the L2 keep the default policy.
template <typename T> __device__ __forceinline__ T ldnt(gptr<const T> p)
{
    return __builtin_nontemporal_load(p);
}
template <typename T> __device__ __forceinline__ T ldnt(const T *p)
{
    return __builtin_nontemporal_load(p);
}

// Stores of outputs that no later work of the same launch reads again, with
// the nontemporal hint.
template <typename T> __device__ __forceinline__ void stnt(gptr<T> p, T v)
{
    __builtin_nontemporal_store(v, p);
}

// 1/x and 1/sqrt(x) in FP64 to about one ulp: the hardware estimates refined
// by two Newton-Raphson steps each (r <- r + r (1 - x r);
// y <- y (3 - x y^2) / 2), each step squaring the relative error.
__device__ __forceinline__ double rcp64(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}
__device__ __forceinline__ double rsq64(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    y = y * fma(-0.5 * x, y * y, 1.5);
    y = y * fma(-0.5 * x, y * y, 1.5);
    return y;
}

}  // namespace pisces
