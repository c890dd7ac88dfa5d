// Stands in for the HIP runtime header when tests/lanes builds the device
// code for the host (see ../hip_lanes.h).
#pragma once
#include "../hip_lanes.h"
