/* Kerr-Schild GR EZ oracle (TEST INFRASTRUCTURE ONLY; see ffe_oracle.h). */
#include "ffe_oracle.h"
