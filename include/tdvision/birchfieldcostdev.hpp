// Reference header LibThunderVision/src/tdvision/birchfieldcostdev.hpp: its declarations (BirchfieldCostDev)
// live in stages.hpp.
#ifndef TDV_FWD_BIRCHFIELDCOSTDEV_HPP
#define TDV_FWD_BIRCHFIELDCOSTDEV_HPP
#include "stages.hpp"
#endif
