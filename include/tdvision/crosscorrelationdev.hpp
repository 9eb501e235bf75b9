// Reference header LibThunderVision/src/tdvision/crosscorrelationdev.hpp: its declarations (CrossCorrelationDev)
// live in stages.hpp.
#ifndef TDV_FWD_CROSSCORRELATIONDEV_HPP
#define TDV_FWD_CROSSCORRELATIONDEV_HPP
#include "stages.hpp"
#endif
