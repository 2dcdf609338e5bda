// noalloc.hip -- the static device scratch of the allocation-free entry points.
//
// gca_build_gauss_newton_noalloc (icp.hip) runs the Gauss-Newton build of
// gca_build_gauss_newton for callers that cannot allocate (code under HIP graph capture, serving
// loops): its workgroup partials and last-block ticket live in `scratch`, a file-scope device
// array of fixed size that every device holds until the process ends.  Host code never passes the
// array's name to a kernel: in host code it is a stand-in, not a device address.  The address is
// resolved once per device, with that device current, by hipGetSymbolAddress, and kept per device;
// kernels receive that address.  Every call on a device shares its array: calls that overlap on
// one device (two streams, two threads) race, and the ticket is cleared at the start of each call.
#include <hip/hip_runtime.h>

#include <mutex>

#include "gca_internal.hpp"

namespace gca {

static_assert(sizeof(NoAllocScratch) <= 264 * 1024, "the no-alloc scratch stays a modest, stated size");

static __device__ __align__(256) unsigned char scratch[sizeof(NoAllocScratch)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_addr[kMaxDevices] = {};
}  // namespace

void *noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_addr[device]) {
        // the caller (gca_ctx_create) has made `device` current
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        g_addr[device] = p;
    }
    return g_addr[device];
}

}  // namespace gca
