/* Shared oracle helpers (TEST INFRASTRUCTURE ONLY; see ffe_oracle.h). */
#include "ffe_oracle.h"

int ffe_oracle_version(void) { return 1; }
