// Declare.hpp -- the header name the reference's sources include
// (src/Declare.hpp) for Config_VVM, class vvm and its operators.  Code
// written against the reference (its own src/main.cpp and src/Iteration.cpp,
// for two) compiles unchanged against this include directory and links with
// libvvm.so; the declarations are those of vvm.hpp, whose operators run on
// the GPU.
#pragma once

// The configuration the device timestep implements, as the reference's
// src/Config.hpp flags name it: Adams-Bashforth time stepping, moisture,
// Kessler microphysics.  TROPICALFORCING is a run-time choice here
// (Config_VVM::tropical_forcing).
#ifndef AB2
#define AB2
#endif
#ifndef WATER
#define WATER
#endif
#ifndef KESSLER_MICROPHY
#define KESSLER_MICROPHY
#endif

// Standard headers the reference's sources rely on through theirs.
#include <algorithm>
#include <cmath>
#include <iostream>
#include <string>
#include <vector>

#include "vvm.hpp"
