// mg.hip -- p-multigrid preconditioner.
#include "mg.hpp"

namespace nrs {

void mg_setup(MGSolver &, nrs_elliptic *, const nrs_elliptic_desc *, hipStream_t)
{
  throw Error(NRS_ERR_UNSUPPORTED, "multigrid preconditioner not built yet");
}

void mg_apply(MGSolver &, const double *, double *, hipStream_t)
{
  throw Error(NRS_ERR_UNSUPPORTED, "multigrid preconditioner not built yet");
}

}  // namespace nrs
