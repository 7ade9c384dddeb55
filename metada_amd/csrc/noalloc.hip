// Allocation-free ETKF entry point (mda_etkf_analyse_noalloc).
//
// For callers that cannot allocate device memory (code under hipGraph
// capture, tight serving loops): the same kernels as mda_etkf_analyse, with
// every buffer the analysis needs beyond its inputs and outputs carved out of
// one file-scope device array, `scratch`.  Each GPU holds its own copy of the
// array from module load until the process ends.  Host code never uses the
// array's name as an address: on the first call on a device (with that device
// current) hipGetSymbolAddress gives the device address, which is kept per
// device and used for every kernel argument.
//
// Every call on a GPU shares that GPU's array and leaves its contents behind;
// the status word is cleared by a kernel at the start of each call and every
// other buffer is written before it is read.  Calls that overlap on one GPU
// (several host threads or contexts) race on the array.
#include <mutex>

#include "mda_internal.hpp"

namespace {

constexpr size_t NOALLOC_DOUBLES = (size_t)13 << 18;  // 3,407,872 doubles = 26 MiB
constexpr int NOALLOC_MAX_DEVICES = 64;

__device__ __align__(256) double scratch[NOALLOC_DOUBLES];

std::mutex g_scratch_mu;
double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// device address of `scratch` on device `dev`, which must be current
double* scratch_address(int dev) {
  MDA_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    MDA_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    MDA_REQUIRE(ptr != nullptr, "hipGetSymbolAddress returned no address");
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  return g_scratch_ptr[dev];
}

size_t align32(size_t n) { return (n + 31) & ~(size_t)31; }  // 256-byte pieces

}  // namespace

extern "C" int mda_etkf_analyse_noalloc(mda_ctx* c, double inflation, mda_ensemble* ens,
                                        const mda_obs* obs, const mda_hop* hop) {
  return mda::guarded_call(c, [&] {
    MDA_REQUIRE(ens && obs && hop, "ensemble, observations and H are required");
    MDA_REQUIRE(ens->on_device && obs->on_device && hop->on_device,
                "mda_etkf_analyse_noalloc takes a device-resident ensemble, observations and H");
    const int k = ens->k;
    const int64_t p = obs->p;
    MDA_REQUIRE(k >= 2 && k <= MDA_NOALLOC_MAX_K,
                "mda_etkf_analyse_noalloc supports 2 <= k <= 64");
    MDA_REQUIRE(p >= 1 && p * (k + 1) <= MDA_NOALLOC_MAX_OBS_VALUES,
                "mda_etkf_analyse_noalloc: p * (k + 1) must be <= 2,097,152 (scratch limit)");
    double* base = scratch_address(c->device);
    size_t off = 0;
    mda::etkf_bufs b;
    b.yp = base + off; off += align32((size_t)p * k);
    b.dd = base + off; off += align32((size_t)p);
    b.part = base + off; off += align32((size_t)mda::ETKF_GG_BLOCKS * k * (k + 1));
    b.EB = base + off; off += align32((size_t)k * (k + 1));
    b.T = base + off; off += align32((size_t)k * k);
    b.scr = base + off; off += align32((size_t)2 * k * k);
    b.status = reinterpret_cast<int*>(base + off); off += 32;
    MDA_REQUIRE(off <= NOALLOC_DOUBLES, "mda_etkf_analyse_noalloc: scratch layout too large");
    mda::etkf_run(c, inflation, ens, obs, hop, &b);
  });
}
