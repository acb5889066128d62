// Instantiates the device BFS and segment-step kernels for the packed puzzle family.
#include "bfs_launch.h"
#include "fs_launch.h"
#include "seg3_launch.h"

namespace ps {
PS_DEFINE_BFS_FAMILY(packed, FAM_PACKED)
PS_DEFINE_FS_FAMILY(packed, FAM_PACKED)
PS_DEFINE_S3_FAMILY(packed, FAM_PACKED)
}  // namespace ps
