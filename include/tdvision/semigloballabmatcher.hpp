// Reference header LibThunderVision/src/tdvision/semigloballabmatcher.hpp: its declarations (SemiglobalLabMatcher)
// live in matchers.hpp.
#ifndef TDV_FWD_SEMIGLOBALLABMATCHER_HPP
#define TDV_FWD_SEMIGLOBALLABMATCHER_HPP
#include "matchers.hpp"
#endif
