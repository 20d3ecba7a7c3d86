// First-pass scan: the table of tile configurations and the launch dispatch.
// The kernels are the templates of scan_impl.h, instantiated per group of
// configurations in kernels_scan_{a,b,c,d}.hip.
#include "launch.h"

namespace mann {

bool launch_scan_group_a(const ScanArgs& a, int mode, int grid, int kind, hipStream_t st);
bool launch_scan_group_b(const ScanArgs& a, int mode, int grid, int kind, hipStream_t st);
bool launch_scan_group_c(const ScanArgs& a, int mode, int grid, int kind, hipStream_t st);
bool launch_scan_group_d(const ScanArgs& a, int mode, int grid, int kind, hipStream_t st);

// Tile configurations: WR x WC wavefronts of 32*WM rows x 32*WN queries.
// Configs 6 and 7 repeat 5 and 0 with a 4-deep LDS ring (3 chunks in flight);
// configs 8 (128 rows x 64 queries) and 9 (128 x 32 on two wavefronts) suit
// items with few queries.
struct ScanCfg {
  int wm, wn, wr, wc, stages;
};
static const ScanCfg kCfg[] = {{2, 2, 2, 2, 3}, {4, 2, 2, 2, 3}, {2, 4, 2, 2, 3}, {4, 4, 2, 2, 3},
                               {4, 2, 2, 4, 3}, {2, 4, 4, 2, 3}, {2, 4, 4, 2, 4}, {2, 2, 2, 2, 4},
                               {2, 1, 2, 2, 3}, {2, 1, 2, 1, 3}};
constexpr int kNumCfg = sizeof(kCfg) / sizeof(kCfg[0]);

int scan_num_cfg() { return kNumCfg; }
int scan_tile_rows(int cfg) {
  const ScanCfg& c = kCfg[cfg];
  return 32 * c.wm * c.wr;
}
int scan_tile_queries(int cfg) {
  const ScanCfg& c = kCfg[cfg];
  return 32 * c.wn * c.wc;
}

void launch_scan(const ScanArgs& a, int mode, int grid, int kind, hipStream_t st) {
  if (grid <= 0) return;
  // row norms are read through a 32-bit byte range
  MANN_REQUIRE(((int64_t)a.n_rows + 1024) * 4 < ((int64_t)1 << 32), ERR_UNSUPPORTED,
               "too many rows for one first-pass launch");
  // each group of tile configurations is its own translation unit
  if (launch_scan_group_a(a, mode, grid, kind, st) || launch_scan_group_b(a, mode, grid, kind, st) ||
      launch_scan_group_c(a, mode, grid, kind, st) || launch_scan_group_d(a, mode, grid, kind, st))
    return;
  throw Error{ERR_ARGUMENT, "unknown first-pass tile configuration"};
}

}  // namespace mann
