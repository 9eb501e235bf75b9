// Reference header LibThunderVision/src/tdvision/matchingcost.hpp: its declarations (MatchingCost, AbstractMatchingCost)
// live in stages.hpp.
#ifndef TDV_FWD_MATCHINGCOST_HPP
#define TDV_FWD_MATCHINGCOST_HPP
#include "stages.hpp"
#endif
