// Reference header LibThunderVision/src/tdvision/stereomatcher.hpp: its declarations (StereoMatcher)
// live in matchers.hpp.
#ifndef TDV_FWD_STEREOMATCHER_HPP
#define TDV_FWD_STEREOMATCHER_HPP
#include "matchers.hpp"
#endif
