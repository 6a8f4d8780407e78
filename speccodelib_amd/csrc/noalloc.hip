// scl_dns_spectra_noalloc: the shell spectra and totals of scl_dns_spectrum,
// scl_dns_helicity_spectrum and scl_dns_diagnostics, computed by the same
// kernels, for callers that cannot allocate device memory (code under HIP
// graph capture, serving loops).  The per-workgroup partial sums live in the
// file-scope array `scratch` below instead of the plan's lazily allocated
// buffer; the result goes to a device array the caller owns.
#include <mutex>

#include "diag.hpp"
#include "plan.hpp"

namespace {

// Largest nbins the array holds (the spectra of a 1024-point x-axis).
constexpr int NOALLOC_MAX_BINS = 512;
constexpr int64_t NOALLOC_SCRATCH_DOUBLES = (int64_t)scl::DIAG_BLOCKS * (2 * NOALLOC_MAX_BINS + scl::DIAG_TOT);

}  // namespace

// 512 x 1032 doubles = 4.03 MiB, held on each device the library runs on
// until the process ends.  Host code never uses this name as an address:
// it is a host-side stand-in; the device address comes from
// hipGetSymbolAddress, once per device.
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double* g_scratch[MAX_DEVICES] = {};

// The device address of `scratch` on the current device `dev`, looked up on
// the first call on that device.
int32_t scratch_of(int dev, double** out) {
  if (dev < 0 || dev >= MAX_DEVICES) return scl::set_error("device %d out of range", dev), SCL_EINVAL;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      scl::set_error("hipGetSymbolAddress(scratch) on device %d failed: %s", dev, hipGetErrorString(e));
      return SCL_EHIP;
    }
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return SCL_OK;
}

}  // namespace

extern "C" int32_t scl_dns_spectra_noalloc(scl_dns* d, double* out, int32_t nbins) {
  using namespace scl;
  if (!d || !out) return set_error("null solver or output"), SCL_EINVAL;
  scl_plan3d* p = d->plan;
  if (nbins < 1 || nbins > p->g.mx)
    return set_error("nbins = %d: must be in [1, nx/2 = %d]", nbins, p->g.mx), SCL_EINVAL;
  if (nbins > NOALLOC_MAX_BINS || diag_part_len(nbins) > NOALLOC_SCRATCH_DOUBLES)
    return set_error("nbins = %d exceeds the %d shells the static scratch holds", nbins, NOALLOC_MAX_BINS),
           SCL_EINVAL;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return (void)hipGetLastError(), set_error("no current device"), SCL_EHIP;
  if (dev != p->device)
    return set_error("device %d is current, the solver's plan is on device %d", dev, p->device), SCL_EINVAL;
  double* part = nullptr;
  SCL_CHECK(scratch_of(dev, &part));
  if (p->prec == SCL_SP)
    return diag_launch<float>(p->g, (const cpx<float>*)d->fields, p->fstride(), nbins, part, out, p->stream);
  return diag_launch<double>(p->g, (const cpx<double>*)d->fields, p->fstride(), nbins, part, out, p->stream);
}
