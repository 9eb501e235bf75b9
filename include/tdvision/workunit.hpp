// Reference header LibThunderVision/src/tdvision/workunit.hpp: its declarations (WorkUnit)
// live in process.hpp.
#ifndef TDV_FWD_WORKUNIT_HPP
#define TDV_FWD_WORKUNIT_HPP
#include "process.hpp"
#endif
