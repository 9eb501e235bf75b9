// Reference header LibThunderVision/src/tdvision/devstereomatcher.hpp: its declarations (DevStereoMatcher)
// live in matchers.hpp.
#ifndef TDV_FWD_DEVSTEREOMATCHER_HPP
#define TDV_FWD_DEVSTEREOMATCHER_HPP
#include "matchers.hpp"
#endif
