// k_gb.hip -- FF_GeneralizedBorn_Still on the device (placeholder until the GB passes land).
#include "engine.hpp"

namespace pdgpu {

void Engine::launch_gb(double *) {
  throw EngineError(PDGPU_ERR_SETUP, "GB component not available in this build");
}

void Engine::gb_energies_to_host(const double *, pdgpu_energies &) {}

}  // namespace pdgpu

extern "C" int pdgpu_setup_gb(pdgpu_engine *, const pdgpu_gb_params *, const double *, const int32_t *,
                              const int32_t *, const double *, const int32_t *, const int32_t *, const double *) {
  return PDGPU_ERR_SETUP;
}

extern "C" int pdgpu_get_born_radii(pdgpu_engine *, double *) { return PDGPU_ERR_SETUP; }
