// setup.hpp -- host-side set-up shared by the drop-in objects and the engine.
#pragma once

#include <cstdint>
#include <vector>

#include "fgr_weights.hpp"
#include "tt06_f64.hpp"
#include "tt06_model.hpp"

namespace cdh {

// TT06Dev honours the rational fits and the fast kernels only with
// tolerance 1e-4 and mod (jhTauSmooth) 1 (elec/TT06Dev_Reaction.cc:110-123);
// this implementation provides configuration B: v2 gates, generic non-gates.
void check_configuration_b(double tolerance, int mod, int fastGate, int fastNonGate);

// Validates the fit orders against the kernels' and returns the 468
// coefficients with the 13 tauR numerators multiplied by dt
// (elec/TT06Dev_Reaction.cc:130-136).
std::vector<double> scaled_fit_coefficients(const int *l, const int *m, const double *coef,
                                            double dt);

ParamSet make_param_set(const CellType &ct, double s_kr, double s_cal, double s_na);

// rest state [19][stride] with dVK_i = K_i/c9 + Vm0 (elec/TT06Dev_Reaction.cc:194-199)
std::vector<double> rest_state_soa(const CellType &ct, int64_t n, int64_t stride, double c9);

// bounding box of the first nLocal cells padded by one (elec/DiffusionUtils.cc:19-66)
BoxGeom bounding_box(int nx, int ny, int nz, const int64_t *gid, int64_t nLocal, double dx,
                     double dy, double dz);
// box index of global cell gid, or -1 if outside the box
int64_t box_index(const BoxGeom &g, int nx, int ny, int64_t gid);

}  // namespace cdh
