// sweep_fp32.hip - the FP32-bulk instantiations of the sweep kernels and their
// launches.  This translation unit lets the compiler fuse multiplies and adds
// (FMA) in every floating-point expression, FP32 and FP64 alike: the FP32-bulk mode
// carries FP32 rounding in its face arithmetic anyway, and a fused operation rounds
// once instead of twice.  The FP64 parity kernels are instantiated in
// gelios_sweep.hip with contraction off.
#pragma clang fp contract(fast)
#include "sweep_kernel.h"

namespace gel {

hipError_t launch_sweep_fp32(gel_ctx* c, SweepArgs& a, bool moving)
{
    const int src = c->prm.source_mode;
    hipStream_t s = c->stream;
    if (c->prm.physics == GEL_PHYS_MHD) return launch_ms<MhdPhys, float, float, false, 3>(c, a, s, moving, src);
    return c->prm.flux == GEL_FLUX_HLLC ? launch_ms<GasPhys, float, float, false, GEL_FLUX_HLLC>(c, a, s, moving, src)
                                        : launch_ms<GasPhys, float, float, false, GEL_FLUX_HLL>(c, a, s, moving, src);
}

hipError_t launch_batch_fp32(const gel_params& p, const SweepArgs* d_args, const BatchItem* d_items, int nitems,
                             int workgroups, hipStream_t s, bool moving)
{
    if (p.physics == GEL_PHYS_MHD)
        return batch_launch_s<MhdPhys, float, float, 3>(d_args, d_items, nitems, workgroups, s, p.source_mode, moving);
    if (p.flux == GEL_FLUX_HLLC)
        return batch_launch_s<GasPhys, float, float, GEL_FLUX_HLLC>(d_args, d_items, nitems, workgroups, s,
                                                                    p.source_mode, moving);
    return batch_launch_s<GasPhys, float, float, GEL_FLUX_HLL>(d_args, d_items, nitems, workgroups, s, p.source_mode,
                                                               moving);
}

int resident_batch_fp32(const gel_params& p, bool moving)
{
    const void* kern = moving ? (p.physics == GEL_PHYS_MHD
                                     ? (const void*)sweep_batch_kernel<MhdPhys, float, float, GEL_SRC_NONE, 3, kNW, true>
                                     : (const void*)sweep_batch_kernel<GasPhys, float, float, GEL_SRC_NONE,
                                                                       GEL_FLUX_HLLC, kNW, true>)
                              : (p.physics == GEL_PHYS_MHD
                                     ? (const void*)sweep_batch_kernel<MhdPhys, float, float, GEL_SRC_TABLE, 3, kNW>
                                     : (const void*)sweep_batch_kernel<GasPhys, float, float, GEL_SRC_TABLE,
                                                                       GEL_FLUX_HLLC, kNW>);
    return resident_groups(kern, kNW * kWave);
}

}  // namespace gel
