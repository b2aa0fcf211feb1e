// schedule.hpp -- MULTIGRID SCHEDULE parsing (parseMultigridSchedule).
#pragma once
#include <string>

#include "nrs_hip.h"

namespace nrs {

// fills o->nLevels, o->levels, o->preDegree, o->postDegree; throws Error
// (NRS_ERR_INVALID) with the reasons on an invalid schedule
void parse_mg_schedule(const std::string &schedule, int N, int defaultDegree, nrs_elliptic_opts *o);

}  // namespace nrs
