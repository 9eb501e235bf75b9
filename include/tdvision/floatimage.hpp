// Reference header LibThunderVision/src/tdvision/floatimage.hpp: its declarations (FloatImage)
// live in images.hpp.
#ifndef TDV_FWD_FLOATIMAGE_HPP
#define TDV_FWD_FLOATIMAGE_HPP
#include "images.hpp"
#endif
