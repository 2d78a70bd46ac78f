// adjoint.hip - placeholder until the PBF adjoint lands.
#include "vtp_internal.h"
namespace vtp {
void autotune(Ctx &) {}
}  // namespace vtp
extern "C" vtp_status vtp_pbf_rollout_fwd(vtp_ctx *, uint32_t, uint32_t) { return VTP_EINVAL; }
extern "C" vtp_status vtp_pbf_rollout_bwd(vtp_ctx *, const float *, const float *, float *, float *, float *) { return VTP_EINVAL; }
extern "C" vtp_status vtp_pbf_density_loss(vtp_ctx *, double *, float *) { return VTP_EINVAL; }
extern "C" vtp_status vtp_pbf_set_xsph(vtp_ctx *, float) { return VTP_EINVAL; }
