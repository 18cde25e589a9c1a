// The voxel kernel for the split launch schedule: the NNLS workspace rows beyond an LDS cutoff live in the
// workgroup's global slot. dmi_kernels.hip builds the same kernel with every row in LDS, where those branches compile
// away, together with the other kernels and the host launchers; this unit adds launch_t2map_split and
// t2map_occupancy_split, which launch_t2map and t2map_occupancy call for split schedules.
#define DMI_SPLIT_TU 1
#include "dmi_kernels.hip"
