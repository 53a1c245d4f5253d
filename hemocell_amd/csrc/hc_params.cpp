// hc_params.cpp -- SI -> lattice-unit conversion and model stiffness constants.
//
// Follows hemo::Parameters::lbm_base_parameters (mechanics/constantConversion.cpp:36-59),
// lbm_pipe_parameters / lbm_shear_parameters (:61-90) and the
// CellMechanics::calculate_k* converters (mechanics/cellMechanics.h:50-78,
// mechanics/wbcHighOrderModel.cpp:243-262).
#include <cmath>

#include "hc_internal.h"

extern "C" {

hc_status hc_lbm_base_parameters(hc_units* u) {
  if (!u || !(u->dx > 0) || !(u->nu_p > 0) || !(u->rho_p > 0)) return HC_INVALID_ARG;
  if (u->dt < 0.0) {
    // dt not set: tau = 1 and dt follows from the viscosity
    u->tau = 1.0;
    u->nu_lbm = 1.0 / 3.0 * (u->tau - 0.5);
    u->dt = u->nu_lbm / u->nu_p * (u->dx * u->dx);
  } else {
    u->nu_lbm = u->nu_p * u->dt / (u->dx * u->dx);
    u->tau = 3.0 * u->nu_lbm + 0.5;
  }
  u->dm = u->rho_p * (u->dx * u->dx * u->dx);
  u->df = u->dm * u->dx / (u->dt * u->dt);
  u->f_limit = 50.0 / 1.0e12 / u->df;  // FORCE_LIMIT = 50 pN
  u->kBT_lbm = u->kBT_p / (u->df * u->dx);
  return HC_OK;
}

hc_status hc_model_constants(int model, const hc_material* m, int n_triangles, const hc_units* u,
                             hc_model_params* out) {
  if (!m || !u || !out || n_triangles <= 0 || !(u->dx > 0) || !(u->dt > 0) || !(u->df > 0)) return HC_INVALID_ARG;
  if (model < HC_MODEL_RBC || model > HC_MODEL_WBC) return HC_INVALID_ARG;
  const double plc = 7.5e-9 / u->dx;   // persistence length, lattice units
  const double eqLength = 5e-7 / u->dx;
  const double faces = 1280.0 / n_triangles;
  out->k_link = m->kLink * u->kBT_lbm / plc;
  out->k_bend = m->kBend * u->kBT_lbm / eqLength;
  out->k_volume = m->kVolume * faces * u->kBT_lbm / eqLength;
  out->k_area = m->kArea * faces * u->kBT_lbm / (eqLength);
  out->eta_m = m->eta_m * u->dx / u->dt / u->df;
  out->k_inner_rigid = 0.0;
  out->k_cytoskeleton = 0.0;
  out->core_radius = 0.0;
  out->radius = 0.0;
  if (model == HC_MODEL_WBC) {
    out->k_inner_rigid = m->kInnerRigid / u->df;
    out->k_cytoskeleton = m->kCytoskeleton / u->df;
    out->core_radius = m->coreRadius / u->dx;
    out->radius = m->radius / u->dx;
  }
  return HC_OK;
}

}  // extern "C"
