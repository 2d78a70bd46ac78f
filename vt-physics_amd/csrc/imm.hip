// imm.hip - placeholder until the IMM / IMM-CT kernels land.
#include "vtp_internal.h"
namespace vtp {
void imm_initialize(Ctx &) { throw std::runtime_error("IMM not built yet"); }
void imm_step(Ctx &) { throw std::runtime_error("IMM not built yet"); }
}  // namespace vtp
