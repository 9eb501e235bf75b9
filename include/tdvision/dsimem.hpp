// Reference header LibThunderVision/src/tdvision/dsimem.hpp: its declarations (DSIMem)
// live in images.hpp.
#ifndef TDV_FWD_DSIMEM_HPP
#define TDV_FWD_DSIMEM_HPP
#include "images.hpp"
#endif
