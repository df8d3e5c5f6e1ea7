// device_init.hpp -- the initial state and trap built on the device
// (fgpe_psi_* / fgpe_vext_*) with the choices the Psi and Vext constructors
// and the shipped plugins make on the host.  Selected by the input key
// `device_init true`; without it the host classes build them, with the
// reference's std::mt19937 noise.  The device noise is Philox4x32-10 keyed
// by rndseed (+ component): with rand_pct = crand_pct = 0 both paths give
// the same state.
#ifndef FGPE_HOST_DEVICE_INIT_HPP
#define FGPE_HOST_DEVICE_INIT_HPP

#include "../../../include/fgpe_c.h"
#include "params.hpp"

// device_init present and true, and the state / trap is not read from file
bool deviceInitPsi(const ParamsBase& p);
bool deviceInitVext(const ParamsBase& p);
fgpe_geom geomOf(const ParamsBase& p);
// Psi(p, type) on the device into component comp (type 0: one component),
// normalised, optionally vortex-imprinted; copied to host_psi.
void devicePsi(fgpe_ctx* ctx, const ParamsBase& p, int comp, int type, bool vortex, complex* host_psi);
// Vext(p, type) on the device plus A cos(2 pi nseed x / L) when seed is set;
// committed to component comp and copied to host_vext.
void deviceVext(fgpe_ctx* ctx, const ParamsBase& p, int comp, int type, bool seed, double A, complex* host_vext);

#endif
