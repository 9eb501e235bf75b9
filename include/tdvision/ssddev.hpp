// Reference header LibThunderVision/src/tdvision/ssddev.hpp: its declarations (SSDDev)
// live in stages.hpp.
#ifndef TDV_FWD_SSDDEV_HPP
#define TDV_FWD_SSDDEV_HPP
#include "stages.hpp"
#endif
