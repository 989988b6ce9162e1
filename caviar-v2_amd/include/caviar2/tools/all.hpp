// caviar2/tools/all.hpp -- the system-builder tools available on this path.
#pragma once
#include "caviar2/tools/atom.hpp"
