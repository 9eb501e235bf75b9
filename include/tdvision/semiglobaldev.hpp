// Reference header LibThunderVision/src/tdvision/semiglobaldev.hpp: its declarations (SemiGlobalDev)
// live in stages.hpp.
#ifndef TDV_FWD_SEMIGLOBALDEV_HPP
#define TDV_FWD_SEMIGLOBALDEV_HPP
#include "stages.hpp"
#endif
