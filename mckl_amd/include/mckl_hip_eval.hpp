// mckl_hip_eval.hpp -- device evaluation objects for the samplers.
//
// The reference's evaluation objects run per particle (eval_each) or per
// range of particles (eval_range) on the host (smp/backend_base.hpp:113-183,
// smp/backend_seq.hpp), calling the host libm through math/vmf.hpp.  Here the
// per-particle part is a functor with a __device__ operator()(i), compiled
// into one kernel by eval_each (one lane per particle, grid-stride), and
// DeviceMath gives it log and log1p rounded once (the correctly rounded
// value), the stand-ins for vmf's log and log1p.
//
// A header for translation units compiled by hipcc for gfx950; link
// libmckl_hip.so.  The functor is copied to the kernel by value: it holds
// device pointers and scalars.
#ifndef MCKL_HIP_EVAL_HPP
#define MCKL_HIP_EVAL_HPP

#include <hip/hip_runtime.h>

#include "mckl_hip.hpp"
#include "../csrc/crmath.hpp"

#include <algorithm>
#include <cstddef>

namespace mckl {

/// log and log1p of doubles for device code, each the correctly rounded
/// value.  log: x positive and normal; log1p: x > -1 and finite.
struct DeviceMath {
    const double *log_c = nullptr;
    const double *sin_a = nullptr;
    const double *cos_a = nullptr;

    __device__ double log(double x) const
    {
        return mckl_hip::crm::log_cr(x, log_c);
    }

    __device__ double log1p(double x) const
    {
        return mckl_hip::crm::log1p_cr(x, log_c);
    }
};

/// the tables of the current device (uploaded on first use)
inline DeviceMath device_math()
{
    DeviceMath m;
    hip_check(mckl_hip_device_math(&m.log_c, &m.sin_a, &m.cos_a));
    return m;
}

template <class F>
__global__ __launch_bounds__(256) void k_eval_each(std::size_t n, F f)
{
    for (std::size_t i = (std::size_t)blockIdx.x * blockDim.x + threadIdx.x;
         i < n; i += (std::size_t)gridDim.x * blockDim.x)
        f(i);
}

/// eval_each on the device: f(i) for every i in [0, n) on current_stream()
/// (smp/backend_base.hpp:113-120 calls eval_each once per particle of the
/// range).  Asynchronous; launch errors raise RuntimeAssert.
template <class F>
inline void eval_each(std::size_t n, const F &f)
{
    if (n == 0)
        return;
    const std::size_t groups = std::min<std::size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_eval_each<F>, dim3((unsigned)groups), dim3(256), 0,
        static_cast<hipStream_t>(current_stream()), n, f);
    if (hipGetLastError() != hipSuccess)
        throw RuntimeAssert("eval_each: kernel launch failed");
}

} // namespace mckl

#endif // MCKL_HIP_EVAL_HPP
