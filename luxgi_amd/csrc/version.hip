// version.hip -- library identification.
#include "common.hpp"

extern "C" const char* luxgi_version(void) { return "luxgi-mi355x 0.1 (gfx950)"; }
