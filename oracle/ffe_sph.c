/* Log-spherical 2D EZ oracle (TEST INFRASTRUCTURE ONLY; see ffe_oracle.h). */
#include "ffe_oracle.h"
