// Bulk cold-rain microphysics for gfx950 (cphopt=2): see micro.h.
#include "micro.h"

namespace cress {

int micro_cloudphy(const Grid &, const Consts &, const Opts &, hipStream_t, float, int, float *const *,
                   const float *, const float *, const float *, const float *, float *, float *, float *,
                   const float *, float *, float *, MicroScratch *) {
  return CRESS_EUNSUPPORTED;
}

int micro_satadjst(const Grid &, const Consts &, const Opts &, hipStream_t, float *const *, const float *,
                   float *, float *, float *, MicroScratch *) {
  return CRESS_EUNSUPPORTED;
}

}  // namespace cress
