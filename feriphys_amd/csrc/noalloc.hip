// fp_sim_count_nonfinite_noalloc: fp_sim_count_nonfinite without any device allocation, for callers
// that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  The per-mesh counters live in a file-scope device array, one copy per GPU, whose
// address is looked up once per device and kept; every call on a GPU shares that GPU's copy, so
// calls that overlap on one GPU race (each call clears the counters it uses before counting).
#include <mutex>

#include "sim_internal.hpp"

namespace {

constexpr uint32_t kNoAllocMaxMeshes = 65536;
constexpr int kMaxDevices = 64;

// the counters of one call: one per mesh
static __device__ __align__(256) unsigned long long scratch[kNoAllocMaxMeshes];

std::mutex g_mu;
unsigned long long* g_ptr[kMaxDevices] = {};

__global__ void zero_counts_kernel(unsigned long long* __restrict__ p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// The device address of `scratch` on `device`, looked up once (with that device current) and kept.
int scratch_on(int device, unsigned long long** out) {
  if (device < 0 || device >= kMaxDevices) return fph::fail(FP_ENODEV, "device index beyond the supported range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) {
    int prev = 0;
    HIP_TRY(hipGetDevice(&prev));
    if (prev != device) HIP_TRY(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (prev != device) HIP_TRY(hipSetDevice(prev));
    if (e != hipSuccess) return fph::fail(FP_EHIP, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_ptr[device] = static_cast<unsigned long long*>(p);
  }
  *out = g_ptr[device];
  return FP_OK;
}

}  // namespace

extern "C" {

int fp_sim_count_nonfinite_noalloc(fp_sim* s, uint64_t* counts) {
  if (!s) return fph::fail(FP_EINVAL, "sim is NULL");
  const size_t M = s->meshes.size();
  if (M == 0) return FP_OK;
  if (!counts) return fph::fail(FP_EINVAL, "counts is NULL");
  if (M > kNoAllocMaxMeshes)
    return fph::fail(FP_EINVAL, "fp_sim_count_nonfinite_noalloc: at most 65536 meshes (use fp_sim_count_nonfinite)");
  unsigned long long* d = nullptr;
  if (int e = scratch_on(s->device, &d)) return e;
  zero_counts_kernel<<<dim3((uint32_t)((M + 255) / 256)), dim3(256), 0, s->stream>>>(d, (uint32_t)M);
  HIP_TRY(hipGetLastError());
  fp::StepArgs a = s->args(false);
  HIP_TRY(fp::launch_nonfinite(a.in, s->all_blocks.p, (uint32_t)s->h_blocks.size(), d, s->stream));
  static_assert(sizeof(unsigned long long) == sizeof(uint64_t), "counter width");
  HIP_TRY(hipMemcpyAsync(counts, d, M * sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return FP_OK;
}

}  // extern "C"
