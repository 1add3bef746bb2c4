// decide_dt without allocation, for callers that cannot allocate device
// memory (code under graph capture, tight serving loops, harnesses that
// forbid hipMalloc).
//
// decide_dt reduces the advective limit through a partial-minimum buffer
// and a device dt slot that the library allocates with the domain's device
// context.  decide_dt_noalloc runs the same two kernels on a file-scope
// device array instead, `scratch`, of kScratchDoubles doubles (66 KB): the
// partial minima, then the dt slot.  Each GPU holds its own copy of the array
// until the process ends; its device address is looked up once per device
// (hipGetSymbolAddress, at the library's set-up for the current device,
// otherwise at the first call on another device) and only that address is
// handed to kernels and copies.
//
// Nothing in `scratch` needs clearing: the final kernel reads exactly the
// partial slots that the partial kernel of the same call writes.  Every call
// on a GPU shares that GPU's array and the library stream, so calls that
// overlap on one GPU (from several host threads) race.
#include <cfloat>

#include "decide_dt.h"
#include "sbfs.h"
#include "sbfs_internal.hpp"

namespace sbfs {

// partial minima (kRedSlots - 64 at most, the bound decide_dt uses), then
// the dt slot
constexpr int kScratchPartials = kRedSlots - 64;
constexpr int kScratchDoubles = kScratchPartials + 64;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device, looked up once per
// device
static int scratch_ptr(double** out) {
  int dev = 0;
  SBFS_HIP(hipGetDevice(&dev));
  SBFS_REQUIRE(dev >= 0 && dev < kMaxDevices, "decide_dt_noalloc: device %d out of range", dev);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    SBFS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SBFS_REQUIRE(p != nullptr, "decide_dt_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

int noalloc_setup() {
  double* p = nullptr;
  return scratch_ptr(&p);
}

}  // namespace sbfs

using namespace sbfs;

extern "C" {

int decide_dt_noalloc(const domain_t* domain, const fluid_t* fluid, const interface_t* interface,
                      double* dt) {
  SBFS_REQUIRE(domain && fluid && interface && dt, "decide_dt_noalloc: null argument");
  DomainCtx* c = domain_ctx_find(domain);
  SBFS_REQUIRE(c != nullptr,
               "decide_dt_noalloc: the domain has no device context (create it first with fluid_init, "
               "interface_init, sbfs_domain_create or decide_dt)");
  if (check_array(domain, &fluid->ux, NA_UX, 8, "ux")) return 1;
  if (check_array(domain, &fluid->uy, NA_UY, 8, "uy")) return 1;
  int rows = 0, nblk = 0;
  if (dt_grid((int)c->nx, (int)c->ny, kScratchPartials, &rows, &nblk)) return 1;
  double* s = nullptr;
  if (scratch_ptr(&s)) return 1;
  if (dt_launch(domain, c, fluid, interface, kScratchPartials, s, s + kScratchPartials)) return 1;
  SBFS_HIP(hipMemcpyAsync(c->host_pin, s + kScratchPartials, sizeof(double), hipMemcpyDeviceToHost, stream()));
  SBFS_HIP(hipStreamSynchronize(stream()));
  *dt = c->host_pin[0];
  return 0;
}

}  // extern "C"
