// Allocation-free HODLR matvec: the fused matvec kernels (matvec.hip) with
// their partial products in a fixed device array instead of buffers the
// handle allocates.
//
// `scratch` is one file-scope device array per GPU, held until the process
// ends.  Host code never passes its name to a kernel (in host code the name
// is a host-side stand-in, not a device address): the device address is
// looked up once per device with hipGetSymbolAddress, with that device
// current, and every launch takes that address.  Every call on one GPU
// shares the array, so calls that overlap on one GPU race; the kernels write
// every partial they later read, so nothing needs clearing between calls.
#include <mutex>

#include "hodlr.hpp"

namespace hmat {
namespace {

// 4 Mi doubles (32 MiB): the fused matvec of the 2^18-row, leaf-256,
// rank-48 operator needs 2.36 Mi doubles at four columns per pass.
constexpr size_t kScratchDoubles = (size_t)4 << 20;
__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (looked up once per device).
double* scratch_on(int device) {
  HMAT_REQUIRE(device >= 0 && device < kMaxDevices, HMAT_ERR_DEVICE, "device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) {
    int cur = 0;
    HIP_CHECK(hipGetDevice(&cur));
    if (cur != device) HIP_CHECK(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) HIP_CHECK(hipSetDevice(cur));
    HIP_CHECK(e);
    HMAT_REQUIRE(p != nullptr, HMAT_ERR_DEVICE, "no device address for the matvec scratch");
    g_ptr[device] = static_cast<double*>(p);
  }
  return g_ptr[device];
}

}  // namespace

void hodlr_matvec_noalloc(HodlrEngine& e, const double* dX, int64_t ldx, int nrhs, double* dY, int64_t ldy) {
  e.matvec_with_scratch(dX, ldx, nrhs, dY, ldy, scratch_on(e.device), kScratchDoubles);
}

size_t hodlr_matvec_noalloc_capacity() { return kScratchDoubles; }

}  // namespace hmat
