// Team-of-lanes march (column resident in registers): placeholder entry
// points until the kernel lands; every call reports that the variant is
// unavailable.
#include <hip/hip_runtime.h>

#include "march_common.h"

int team_launch(const MarchArgs &, int, int, int, int, int, hipStream_t) { return -1; }
int team_field_in(const MarchArgs &, int, hipStream_t) { return -1; }
bool team_shape_supported(int, int, int) { return false; }
void team_auto_shape(int, int, int *rows, int *lanes) {
  *rows = 0;
  *lanes = 0;
}
