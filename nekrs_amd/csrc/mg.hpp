// mg.hpp -- p-multigrid preconditioner (MGSolver_t / pMGLevel analogue).
#pragma once
#include "elliptic.hpp"

namespace nrs {

struct MGSolver {
  int placeholder = 0;
};

void mg_setup(MGSolver &mg, nrs_elliptic *el, const nrs_elliptic_desc *d, hipStream_t s);
// z = V-cycle(r), FP64 in/out (casts around the FP32 cycle as ellipticPreconditioner.cpp:58-61)
void mg_apply(MGSolver &mg, const double *r, double *z, hipStream_t s);

}  // namespace nrs
