// multifluid.hip - the plasma + four hydrogen fluids sweep (Setka::Go_5_komponent_2,
// Setka.cpp:8951-9777): the gas sweep instantiated with kMfFlux and launched once per
// fluid on the context's stream (FP64, FP contraction off as every FP64 kernel), after
// one launch that computes the charge-exchange sources of all five fluids from the
// step's 21-field state.  Each sweep launch reads its fluid's arrays and sources and
// writes its fluid's arrays of the next state; the launches share the step's time-step
// slot (an integer atomicMin, order-independent).
#include "sweep_kernel.h"

namespace gel {

template <bool TYPES, int FX>
static hipError_t mf_sweep(gel_ctx* c, SweepArgs& a, hipStream_t s, bool moving)
{
    if (moving) return launch_t<GasPhys, double, double, true, TYPES, GEL_SRC_NONE, FX | kMfFlux>(c, a, s);
    return launch_t<GasPhys, double, double, false, TYPES, GEL_SRC_NONE, FX | kMfFlux>(c, a, s);
}

hipError_t launch_mf_step(gel_ctx* c, SweepArgs& a, bool moving)
{
    const bool hllc = c->prm.flux == GEL_FLUX_HLLC;
    // the charge-exchange sources of all five fluids, once per step
    if (!c->d_mfq && hipMalloc((void**)&c->d_mfq, 20 * c->nc * sizeof(double)) != hipSuccess) return hipErrorOutOfMemory;
    a.mfq = c->d_mfq;
    const dim3 g((unsigned)((c->nc + 255) / 256));
    if (c->has_types) hipLaunchKernelGGL(mf_source_kernel<true>, g, dim3(256), 0, c->stream, a, c->d_mfq);
    else hipLaunchKernelGGL(mf_source_kernel<false>, g, dim3(256), 0, c->stream, a, c->d_mfq);
    for (int f = 0; f <= 4; f++) {
        a.fluid = f;
        // the plasma: ro, p, u, v, Q; fluid Hf: ro_Hf, p_Hf, u_Hf, v_Hf and the plasma Q
        const int o = f == 0 ? 0 : 5 + 4 * (f - 1);
        for (int k = 0; k < 4; k++) a.voff[k] = o + k;
        a.voff[4] = 4;
        hipError_t e;
        if (c->has_types)
            e = hllc ? mf_sweep<true, GEL_FLUX_HLLC>(c, a, c->stream, moving) : mf_sweep<true, GEL_FLUX_HLL>(c, a, c->stream, moving);
        else
            e = hllc ? mf_sweep<false, GEL_FLUX_HLLC>(c, a, c->stream, moving) : mf_sweep<false, GEL_FLUX_HLL>(c, a, c->stream, moving);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace gel
