// FDTDinit.hpp -- the header the reference's FD driver includes
// (src/fd/FDTDmain.cpp:7), here over the drop-in classes of FDTD.hpp.  It
// also brings the names the reference's headers put in scope for their
// callers (FDTDutils.hpp:7 `using namespace std`, utils.hpp:6
// `using namespace std::chrono`), so a driver written against them compiles.
#pragma once

#include <chrono>
#include <iostream>
#include <utility>
#include <vector>

#include "FDTD.hpp"

using namespace std;
using namespace std::chrono;
