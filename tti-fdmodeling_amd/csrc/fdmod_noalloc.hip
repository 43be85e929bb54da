// fdmod_run_energy_noalloc: the time loop with the energy diagnostic (fdmod_run after
// fdmod_set_energy(ctx, 1), BT:926-952) for callers that cannot allocate device memory
// -- code under HIP graph capture, serving loops that must not call hipMalloc.  Its
// device scratch (the energy kernels' per-block partial sums and the energy series of
// the call) is a file-scope array, one copy per GPU that the library runs on, held
// until the process ends; nothing is allocated or freed per call.
#include <algorithm>
#include <mutex>

#include "fdmod_ctx.h"

using namespace fdm;

namespace {

constexpr int kEnergyBlocks = 512;  // as fdmod_set_energy: same partial-sum grouping, same results
constexpr int kNoAllocMaxSteps = FDMOD_NOALLOC_MAX_STEPS;
// [2 * kEnergyBlocks] partial sums, then E, Ek, Ep of up to kNoAllocMaxSteps iterations
constexpr size_t kScratchFloats = 2 * kEnergyBlocks + 3 * (size_t)kNoAllocMaxSteps;

__device__ __align__(256) float scratch[kScratchFloats];

// The array's device address on each GPU, looked up once per device with that device
// current.  (On the host, the name `scratch` is a stand-in, not a device address.)
constexpr int kMaxDevices = 64;
std::mutex g_lock;
float *g_ptr[kMaxDevices] = {};

int scratch_ptr(fdmod_ctx *c, float **out)
{
  int dev = 0;
  HIPCALL(c, hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(c, FDMOD_EDEVICE, "device %d out of range", dev);
  std::lock_guard<std::mutex> hold(g_lock);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(c, FDMOD_EDEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_ptr[dev] = (float *)p;
  }
  *out = g_ptr[dev];
  return FDMOD_OK;
}

}  // namespace

extern "C" int fdmod_run_energy_noalloc(fdmod_ctx *c, int k0, int count, float *E, float *Ek, float *Ep,
                                        double *seconds)
{
  if (!c) return FDMOD_EINVAL;
  if (!c->model_set) return fail(c, FDMOD_ESTATE, "fdmod_run_energy_noalloc before fdmod_set_model");
  if (k0 < 0 || count < 0) return fail(c, FDMOD_EINVAL, "bad iteration range");
  if (count > 0 && (!E || !Ek || !Ep)) return fail(c, FDMOD_EINVAL, "energy output arrays are NULL");
  if (count > kNoAllocMaxSteps)
    return fail(c, FDMOD_EINVAL, "fdmod_run_energy_noalloc runs at most %d iterations per call (asked %d)",
                kNoAllocMaxSteps, count);
  if (c->wsave > 0)  // the snapshot sink stages fields through host memory the library allocates
    return fail(c, FDMOD_EINVAL, "fdmod_run_energy_noalloc does not write snapshots (wsave = %d)", c->wsave);
  float *sp = nullptr;
  int rc = scratch_ptr(c, &sp);
  if (rc) return rc;
  float *partial = sp, *dE = sp + 2 * kEnergyBlocks, *dEk = dE + kNoAllocMaxSteps, *dEp = dEk + kNoAllocMaxSteps;
  const fdmod_params &p = c->p;
  const int k1 = std::min(k0 + count, p.nt);
  const int n = std::max(k1 - k0, 0);
  SweepArgs sa = sweep_args(c);
  const int energy = c->energy;  // the context's own series (if any) is not advanced
  c->energy = 0;
  HIPCALL(c, hipEventRecord(c->ev0, c->stream));
  for (int k = k0; k < k1 && rc == FDMOD_OK; k++) {
    const int it = p.mode == 1 ? k : p.nt - 1 - k;
    rc = step(c, it, sa);
    if (rc == FDMOD_OK) {
      launch_energy_partial(c->stream, c->g, p.nx, p.nz, c->d, partial, kEnergyBlocks);
      launch_energy_final(c->stream, partial, kEnergyBlocks, p.dx, p.dz, dE, dEk, dEp, k - k0);
    }
  }
  c->energy = energy;
  if (rc) return rc;
  HIPCALL(c, hipEventRecord(c->ev1, c->stream));
  if (n > 0) {
    HIPCALL(c, hipMemcpyAsync(E, dE, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCALL(c, hipMemcpyAsync(Ek, dEk, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCALL(c, hipMemcpyAsync(Ep, dEp, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  }
  HIPCALL(c, hipStreamSynchronize(c->stream));
  HIPCALL(c, hipGetLastError());
  float ms = 0.0f;
  HIPCALL(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  if (seconds) *seconds = ms * 1e-3;
  return FDMOD_OK;
}
