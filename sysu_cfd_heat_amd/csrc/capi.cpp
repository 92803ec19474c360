// Error channel and parameter defaults of the C-ABI.
#include <cstring>
#include <string>

#include "cfd_solver.h"

namespace cfd {
std::string &last_error() {
    static thread_local std::string e;
    return e;
}
}  // namespace cfd

extern "C" {

const char *cfd_last_error(void) { return cfd::last_error().c_str(); }

void cfd_params_default(cfd_params *p) {
    std::memset(p, 0, sizeof(*p));
    // NavierStokesSolver::NavierStokesSolver (src/navier.cpp:354-416)
    p->steady = 1;
    p->max_step = 10000;
    p->residual_steady = 1.e-6;
    p->solve_energy = 0;
    p->turb_model = 0;
    p->density_model = 0;
    p->cht = 0;
    p->pressure_ref = 1.01325e5;
    p->cell_pressure_ref = 0;
    p->urf[0] = p->urf[1] = p->urf[2] = 0.6;
    p->urf[3] = 0.5;
    p->urf[4] = p->urf[5] = p->urf[6] = p->urf[7] = 0.8;
    p->limiter = 0;
    p->prl = 0.72;
    p->prte = 0.9;
    p->cp = 1006.;
    // solver settings at the call sites
    p->mom = cfd_solve_cfg{CFD_GMRES, CFD_PC_SSOR, 1.2, 1.e-8, 500};   // navier_velocity.cpp:26
    p->pres = cfd_solve_cfg{CFD_CG, CFD_PC_SSOR, 1.3, 1.e-6, 1000};   // navier_pressure.cpp:29
    p->ke = cfd_solve_cfg{CFD_GMRES, CFD_PC_SSOR, 1.3, 1.e-8, 500};    // navier_scalar.cpp:101,143
    p->energy = cfd_solve_cfg{CFD_GMRES, CFD_PC_SSOR, 1.3, 1.e-8, 500}; // navier_scalar.cpp:374
    p->solid = cfd_solve_cfg{CFD_CG, CFD_PC_SSOR, 1.3, 1.e-8, 500};    // HeatConductionSolver.cpp:90
    p->solid_heat_source = 0.;
    p->time_scheme = 1;      // CycasSolver::TimeScheme (src/BasicSolver.cpp:10)
    p->dt = 0.;              // src/BasicSolver.cpp:9
    p->total_time = 0.;      // src/navier.cpp:371
    p->max_outer_step = 50;  // src/navier.cpp:367
    p->pres_precision = 0;
    p->multi_rhs = 0;
    p->pres_project = 0;
}

}  // extern "C"
