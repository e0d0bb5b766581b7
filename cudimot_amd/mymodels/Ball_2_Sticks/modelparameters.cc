#include "modelparameters.h"
// Rows of each common fixed parameter (per measurement) and of each
// per-voxel fixed parameter.
int MODEL::CFP_size[] = {3, 1};
int MODEL::FixP_size[] = {};
