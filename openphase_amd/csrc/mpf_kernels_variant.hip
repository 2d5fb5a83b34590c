// One variant of the per-cell step and deferred-cell kernels (mpf_step_kernels.h), compiled once
// per precision and variant with MPF_VARIANT_FP32 (0 FP64, 1 FP32) and MPF_VARIANT_FULL (0 plain
// isotropic, 1 orientation-dependent energies and driving force), so the four variants build in
// parallel. The launchers sit in namespace mpf::v{64,32}_{plain,full} (mpf_device.h).
#include "mpf_step_kernels.h"

#if MPF_VARIANT_FP32
#define MPF_TV float
#if MPF_VARIANT_FULL
#define MPF_VNS v32_full
#else
#define MPF_VNS v32_plain
#endif
#else
#define MPF_TV double
#if MPF_VARIANT_FULL
#define MPF_VNS v64_full
#else
#define MPF_VNS v64_plain
#endif
#endif

namespace mpf {
namespace MPF_VNS {

int launch_step(const StepArgs& one, const StepArgs* args, const uint32_t* first, int ndom, unsigned blocks, bool uni,
                void* stream)
{
    const dim3 grid(blocks);
    hipStream_t s = static_cast<hipStream_t>(stream);
#if MPF_VARIANT_FULL
    (void)uni;
    hipLaunchKernelGGL((step_kernel<MPF_TV, MPF_TV, true, true, false>), grid, dim3(kBlock), 0, s, one, args, first, ndom);
#else
    if (uni) hipLaunchKernelGGL((step_kernel<MPF_TV, MPF_TV, false, false, true>), grid, dim3(kBlock), 0, s, one, args, first, ndom);
    else     hipLaunchKernelGGL((step_kernel<MPF_TV, MPF_TV, false, false, false>), grid, dim3(kBlock), 0, s, one, args, first, ndom);
#endif
    return static_cast<int>(hipGetLastError());
}

int launch_deferred(const StepArgs* args, int ndom, int per_domain, void* stream)
{
    constexpr bool F = MPF_VARIANT_FULL != 0;
    const dim3 grid(static_cast<unsigned>(per_domain) * static_cast<unsigned>(ndom));
    const dim3 grid2(2u * static_cast<unsigned>(per_domain) * static_cast<unsigned>(ndom));
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL((mid_kernel<MPF_TV, MPF_TV, F, F, 0>), grid, dim3(kWideBlock), 0, s, args, ndom);
    hipLaunchKernelGGL((mid_tail_kernel<MPF_TV, MPF_TV, F, F>), grid2, dim3(kWideBlock), 0, s, args, ndom);
    hipLaunchKernelGGL((wide_kernel<MPF_TV, MPF_TV, F, F>), grid, dim3(kWideBlock), 0, s, args, ndom);
    return static_cast<int>(hipGetLastError());
}

}  // namespace MPF_VNS
}  // namespace mpf
