// gkm_config.hpp -- validation of a gkm_config and the gas and grid it implies,
// shared by gkm_create and the entry points that take a configuration.
#pragma once
#include <cmath>
#include <string>

#include "../../include/dnsgkm.h"
#include "gkm_math.hpp"

namespace gkm {

// empty when cfg is usable; otherwise what is wrong with it
inline std::string config_error(const gkm_config& cfg, bool (*grid_ok)(int))
{
    if (cfg.nc > 1024 || !grid_ok(cfg.nc)) return "nc must be in [16, 1024]";
    if (!(cfg.gam > 1.0) || !(cfg.p0 > 0.0) || !(cfg.rho0 > 0.0) || !(cfg.mu0 >= 0.0))
        return "gam must exceed 1 and p0, rho0 must be positive";
    if (!(cfg.cfl > 0.0) || !(cfg.prandtl > 0.0)) return "cfl and prandtl must be positive";
    if (cfg.precision != GKM_FP64 && cfg.precision != GKM_MIXED) return "precision must be GKM_FP64 or GKM_MIXED";
    return std::string();
}

// cell width of the 2 pi periodic cube (FlowField.cpp:46)
inline double config_dx(const gkm_config& cfg) { return 2.0 * (std::atan(1.0) * 4.0) / cfg.nc; }

// gas of the run: K = int((5 - 3 gam)/(gam - 1)) (FlowField.cpp:42), R = 287
// (param.h:21), T0 = p0/(R rho0) (FlowField.cpp:47), WENO eps 1e-10 (param.h:18)
inline Gas<double> config_gas(const gkm_config& cfg)
{
    const int K = int((5.0 - 3.0 * cfg.gam) / (cfg.gam - 1.0));
    const double R = 287.0;
    return Gas<double>{cfg.gam, R, 1.0 / cfg.prandtl, cfg.mu0, cfg.p0 / (R * cfg.rho0), 1e-10, K};
}

}  // namespace gkm
