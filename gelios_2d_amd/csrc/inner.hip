// inner.hip - the inner-region solver (Setka::Go_stationary_5_komponent_inner_MK2,
// Setka.cpp:6999-7212): the static gas sweep instantiated with kInnerFlux (FP64, FP
// contraction off as every FP64 kernel).
#include "sweep_kernel.h"

namespace gel {

template <bool TYPES, int FX>
static hipError_t inner_sweep(gel_ctx* c, SweepArgs& a, hipStream_t s)
{
    if (c->prm.source_mode == GEL_SRC_MK1)
        return launch_t<GasPhys, double, double, false, TYPES, GEL_SRC_MK1, FX | kInnerFlux>(c, a, s);
    return launch_t<GasPhys, double, double, false, TYPES, GEL_SRC_NONE, FX | kInnerFlux>(c, a, s);
}

hipError_t launch_inner_step(gel_ctx* c, SweepArgs& a)
{
    const bool hllc = c->prm.flux == GEL_FLUX_HLLC;
    if (c->has_types) return hllc ? inner_sweep<true, GEL_FLUX_HLLC>(c, a, c->stream) : inner_sweep<true, GEL_FLUX_HLL>(c, a, c->stream);
    return hllc ? inner_sweep<false, GEL_FLUX_HLLC>(c, a, c->stream) : inner_sweep<false, GEL_FLUX_HLL>(c, a, c->stream);
}

}  // namespace gel
