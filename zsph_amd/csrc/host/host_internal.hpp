// Host-side helpers shared by the C ABI translation units (not exported).
#pragma once
#include <cmath>
#include <string>

#include "zsph.h"

int zsph_check_dt(const zsph_cpu_param *p);

// Inverse EOS used to set a hydrostatic initial density:
// gamma == 1: rho = rho0 + p / c^2 (the reference's linear EOS, lib/time_int.cu:39-40);
// otherwise Tait: rho = rho0 (1 + p / B)^(1/gamma), B = rho0 c^2 / gamma.
inline double zsph_host_rho_of_p(double p, double rho0, double c2, double gamma) {
    if (gamma == 1.0) return rho0 + p / c2;
    const double B = rho0 * c2 / gamma;
    return rho0 * std::pow(1.0 + p / B, 1.0 / gamma);
}
