// Reference header LibThunderVision/src/tdvision/fastwtamatcher.hpp: its declarations (FastWTAMatcher)
// live in matchers.hpp.
#ifndef TDV_FWD_FASTWTAMATCHER_HPP
#define TDV_FWD_FASTWTAMATCHER_HPP
#include "matchers.hpp"
#endif
