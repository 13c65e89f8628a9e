// pack_info for callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// sn_pack_info needs a workspace (per-ray counts + the scan's tile totals).
// Here that workspace is a file-scope device array, `scratch`, of a fixed
// size: every GPU the library runs on holds its own copy for the life of the
// process.  Its device address is looked up once per device with
// hipGetSymbolAddress (the host-side name is only a stand-in, not a device
// address) and that address is what the kernels receive.  The kernels are the
// ones sn_pack_info launches: they clear the counts with a fill kernel before
// counting and overwrite every tile total they read, so nothing is carried
// over from the previous call.  All calls on one GPU share its array: calls
// that overlap on one GPU (two streams, two host threads) race.
#include <mutex>

#include "sn_common.h"

namespace sn {

// Largest n_rays sn_pack_info_noalloc accepts; the array holds exactly the
// workspace sn_pack_info asks for at that size (16 MiB of counts + 16 KiB of
// tile totals).
constexpr int64_t kNoAllocMaxRays = int64_t(1) << 22;
constexpr int64_t kNoAllocScanTile = 1024;  // scan.hip's tile (256 lanes x 4 items)
constexpr size_t kNoAllocBytes =
    (size_t)((kNoAllocMaxRays * 4 + 255) / 256 * 256) +
    (size_t)((((kNoAllocMaxRays + kNoAllocScanTile - 1) / kNoAllocScanTile + 1) * 4 + 255) /
             256 * 256);

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace sn

static __device__ __align__(256) unsigned char scratch[sn::kNoAllocBytes];

namespace sn {

// Device address of the current device's copy of `scratch` (looked up on the
// first call on that device, with it current).
static int scratch_address(void** out) {
  int dev = -1;
  SN_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("sn_pack_info_noalloc: device %d outside the %d supported", dev, kMaxDevices);
    return SN_EINVAL;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    SN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      set_error("sn_pack_info_noalloc: no device address for the scratch array");
      return SN_EHIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return SN_OK;
}

}  // namespace sn

extern "C" {

int64_t sn_pack_info_noalloc_max_rays(void) { return sn::kNoAllocMaxRays; }

int sn_pack_info_noalloc(const int64_t* ray_indices, int64_t n_samples, int32_t n_rays,
                         int32_t* packed_info, int32_t* n_bad, void* stream) {
  sn::clear_error();
  SN_REQUIRE(n_rays >= 0 && n_samples >= 0, "sn_pack_info_noalloc: negative size");
  SN_REQUIRE(n_rays <= sn::kNoAllocMaxRays,
             "sn_pack_info_noalloc: n_rays %d exceeds the scratch limit of %lld rays "
             "(use sn_pack_info with a workspace)", n_rays, (long long)sn::kNoAllocMaxRays);
  const size_t need = sn_pack_info_workspace_bytes(n_rays);
  if (need > sizeof(scratch)) {
    sn::set_error("sn_pack_info_noalloc: needs %zu bytes of scratch, has %zu", need,
                  sizeof(scratch));
    return SN_ESPACE;
  }
  void* ws = nullptr;
  int rc = sn::scratch_address(&ws);
  if (rc) return rc;
  return sn_pack_info(ray_indices, n_samples, n_rays, packed_info, n_bad, ws, sizeof(scratch),
                      stream);
}

}  // extern "C"
