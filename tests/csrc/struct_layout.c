/* Prints sizeof and offsetof of the C ABI descriptor structs of include/nebula_hip.h, one "name value" per line,
 * for the ctypes layout test. */
#include <stddef.h>
#include <stdio.h>
#include "../../include/nebula_hip.h"

#define OFF(T, f) printf(#T "." #f " %zu\n", offsetof(T, f))

int main(void) {
  printf("NvStageDesc %zu\n", sizeof(NvStageDesc));
  OFF(NvStageDesc, u); OFF(NvStageDesc, xi); OFF(NvStageDesc, u0); OFF(NvStageDesc, out); OFF(NvStageDesc, stars);
  OFF(NvStageDesc, jets); OFF(NvStageDesc, smax_out); OFF(NvStageDesc, dt_dev); OFF(NvStageDesc, nx);
  OFF(NvStageDesc, ny); OFF(NvStageDesc, nz); OFF(NvStageDesc, nstars); OFF(NvStageDesc, wind);
  OFF(NvStageDesc, jet_cos); OFF(NvStageDesc, jet_r); OFF(NvStageDesc, pulse); OFF(NvStageDesc, dt);
  OFF(NvStageDesc, dx); OFF(NvStageDesc, a); OFF(NvStageDesc, b); OFF(NvStageDesc, floor_); OFF(NvStageDesc, cn2);
  OFF(NvStageDesc, ci2); OFF(NvStageDesc, vmax);
  printf("NvRayDesc %zu\n", sizeof(NvRayDesc));
  OFF(NvRayDesc, rho); OFF(NvRayDesc, nx); OFF(NvRayDesc, ny); OFF(NvRayDesc, nz); OFF(NvRayDesc, sx);
  OFF(NvRayDesc, sy); OFF(NvRayDesc, sz); OFF(NvRayDesc, Q); OFF(NvRayDesc, dr); OFF(NvRayDesc, nsteps);
  OFF(NvRayDesc, dx); OFF(NvRayDesc, rcap); OFF(NvRayDesc, bx); OFF(NvRayDesc, by); OFF(NvRayDesc, bz);
  OFF(NvRayDesc, rfront); OFF(NvRayDesc, nu); OFF(NvRayDesc, nphi); OFF(NvRayDesc, wide);
  OFF(NvRayDesc, bricks); OFF(NvRayDesc, dirs); OFF(NvRayDesc, rf_range);
  printf("NvIonDesc %zu\n", sizeof(NvIonDesc));
  OFF(NvIonDesc, rfronts); OFF(NvIonDesc, nu); OFF(NvIonDesc, nphi); OFF(NvIonDesc, star_xyz);
  OFF(NvIonDesc, nstars); OFF(NvIonDesc, dx); OFF(NvIonDesc, width); OFF(NvIonDesc, xi); OFF(NvIonDesc, dfront_h);
  OFF(NvIonDesc, dfront_f); OFF(NvIonDesc, nx); OFF(NvIonDesc, ny); OFF(NvIonDesc, nz); OFF(NvIonDesc, state);
  OFF(NvIonDesc, cn2); OFF(NvIonDesc, ci2); OFF(NvIonDesc, smax_out); OFF(NvIonDesc, rf_ranges);
  printf("NvPackParams %zu\n", sizeof(NvPackParams));
  OFF(NvPackParams, c_ha); OFF(NvPackParams, c_o3); OFF(NvPackParams, c_s2); OFF(NvPackParams, c_dg);
  OFF(NvPackParams, c_sh); OFF(NvPackParams, gs); OFF(NvPackParams, ambient); OFF(NvPackParams, edge);
  OFF(NvPackParams, inv_L); OFF(NvPackParams, dx);
  return 0;
}
