// Reference header LibThunderVision/src/tdvision/commonstereomatcherfactory.hpp: its declarations (CommonStereoMatcherFactory)
// live in matchers.hpp.
#ifndef TDV_FWD_COMMONSTEREOMATCHERFACTORY_HPP
#define TDV_FWD_COMMONSTEREOMATCHERFACTORY_HPP
#include "matchers.hpp"
#endif
