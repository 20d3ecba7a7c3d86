// Instantiations of the first-pass scan for tile configurations 1, 2
// (one translation unit per group of configurations, so that the library's
// build compiles them in parallel).
#include "scan_impl.h"

namespace mann {

bool launch_scan_group_b(const ScanArgs& a, int mode, int grid, int kind, hipStream_t st) {
  switch (a.cfg) {
    case 1: launch_cfg<4, 2, 2, 2, 3>(a, mode, grid, kind, st); return true;
    case 2: launch_cfg<2, 4, 2, 2, 3>(a, mode, grid, kind, st); return true;
    default: return false;
  }
}

}  // namespace mann
