// messages_to_scalar without any allocation, for callers that cannot call
// hipMalloc (code under graph capture, tight serving loops, harnesses that
// forbid allocation).
//
// Same results and call semantics as cubbs_bbs_gpu_messages_to_scalar_batch_v2
// (reference semantics native/cuda/src/cubbs_bbs_cuda.cu:353-447: one DST per
// call, 32-byte big-endian scalars), on the same hash kernel.  Its device
// memory is one file-scope array, `scratch`, of NOALLOC_SCRATCH_BYTES, which
// each GPU holds until the process ends; its host staging is a static buffer
// of the same size.  A call whose packed input and output (job records, DST,
// message bytes, 32 bytes per item, each rounded up to 256 bytes) exceed that
// size fails with code 1 before touching the device.  Calls share the array
// and the staging buffer, so concurrent calls would race on them: the entry
// point serialises them with a lock.  It calls no hipMalloc, hipFree or
// allocating library routine, and does not record the scalars for
// cubbs_bbs_gpu_sign_e_from_resident_scalars_batch_v2 (the allocating entry
// does).
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "../../include/cubbs_bbs_gpu_abi.h"
#include "../../include/cubbs_hip_ext.h"
#include "bbs_internal.h"

namespace cubbs {

constexpr size_t NOALLOC_SCRATCH_BYTES = CUBBS_HIP_NOALLOC_SCRATCH_BYTES;

// the device memory of every call (host code must use scratch_ptr[device],
// never this name: on the host it is only a stand-in for the symbol)
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex noalloc_mu;
// per device: the device address of `scratch` (hipGetSymbolAddress, once per
// device, with that device current); nullptr until the first call there
void* scratch_ptr[MAX_DEVICES] = {};
alignas(256) uint8_t host_stage[NOALLOC_SCRATCH_BYTES];

size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }
bool slice_ok(const CuBbsBbsGpuByteSlice& s) { return s.len == 0 || s.data != nullptr; }

}  // namespace
}  // namespace cubbs

using namespace cubbs;

extern "C" int32_t cubbs_hip_messages_to_scalar_noalloc_batch(uint32_t abi_version,
                                                              CuBbsBbsGpuMessagesToScalarItem* items,
                                                              size_t item_count,
                                                              CuBbsBbsGpuResult* results) {
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION || (item_count && (!items || !results)))
    return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  if (item_count == 0) return CUBBS_BBS_GPU_SUCCESS;
  // the validation of cubbs_bbs_gpu_messages_to_scalar_batch_v2
  size_t total = 0;
  for (size_t i = 0; i < item_count; i++) {
    const auto& it = items[i];
    if (!slice_ok(it.message)) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    if (it.dst.len > 255 || !slice_ok(it.dst) || !it.scalar_out.data || it.scalar_out.capacity < 32)
      return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    total += it.message.len;
    if (total > NOALLOC_SCRATCH_BYTES) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  }
  const auto& d0 = items[0].dst;
  for (size_t i = 1; i < item_count; i++)
    if (items[i].dst.len != d0.len || (d0.len && std::memcmp(items[i].dst.data, d0.data, d0.len) != 0))
      return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  if (item_count > NOALLOC_SCRATCH_BYTES / 48) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  // layout of the scratch array: job records | DST | message bytes | scalars
  const size_t o_jobs = 0;
  const size_t o_dst = up256(o_jobs + sizeof(H2sJob) * item_count);
  const size_t o_blob = up256(o_dst + std::max<size_t>(d0.len, 1));
  const size_t o_out = up256(o_blob + std::max<size_t>(total, 1));
  const size_t need = up256(o_out + 32 * item_count);
  if (need > NOALLOC_SCRATCH_BYTES) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;

  std::lock_guard<std::mutex> lock(noalloc_mu);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return CUBBS_BBS_GPU_ERR_INTERNAL;
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return CUBBS_BBS_GPU_ERR_INTERNAL;
    }
    scratch_ptr[dev] = p;
  }
  uint8_t* d = static_cast<uint8_t*>(scratch_ptr[dev]);
  // pack on the host
  auto* jobs = reinterpret_cast<H2sJob*>(host_stage + o_jobs);
  size_t cur = 0;
  for (size_t i = 0; i < item_count; i++) {
    const size_t len = items[i].message.len;
    jobs[i] = H2sJob{cur, (uint32_t)len, (uint32_t)i};
    if (len) std::memcpy(host_stage + o_blob + cur, items[i].message.data, len);
    cur += len;
  }
  if (d0.len) std::memcpy(host_stage + o_dst, d0.data, d0.len);
  hipStream_t s = nullptr;  // the null stream: no stream to create
  if (hipMemcpyAsync(d, host_stage, o_out, hipMemcpyHostToDevice, s) != hipSuccess) {
    (void)hipGetLastError();
    return CUBBS_BBS_GPU_ERR_INTERNAL;
  }
  launch_hash_jobs(s, 2, reinterpret_cast<const H2sJob*>(d + o_jobs), (uint32_t)item_count, d + o_blob,
                   d + o_dst, (uint32_t)d0.len, d + o_out);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(host_stage + o_out, d + o_out, 32 * item_count, hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    (void)hipGetLastError();
    return CUBBS_BBS_GPU_ERR_INTERNAL;
  }
  for (size_t i = 0; i < item_count; i++) {
    std::memcpy(items[i].scalar_out.data, host_stage + o_out + 32 * i, 32);
    items[i].scalar_out.len = 32;
    results[i].code = CUBBS_BBS_GPU_SUCCESS;
    results[i].value = 0;
    results[i].bytes_written = 32;
  }
  return CUBBS_BBS_GPU_SUCCESS;
}
