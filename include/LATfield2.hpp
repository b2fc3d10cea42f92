/*
 * LATfield2.hpp -- the drop-in entry header.  A driver written for LATfield2 includes
 * "LATfield2.hpp"; with this directory on the include path ahead of the reference's
 * (-I<this repository>/include) and -llf2mi on the link line, it builds against the
 * MI355X-native particle-mesh path instead: the classes, free functions, `parallel` and
 * COUT that LATfield2.hpp:63-115 provides for that path come from LATfield2_mi.hpp.
 */
#ifndef LATFIELD2_HPP
#define LATFIELD2_HPP
#include "LATfield2_mi.hpp"
#endif
