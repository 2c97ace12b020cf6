// Device memory helpers of the C ABI (for callers without their own HIP binding).
#include <string.h>

#include "hkg_engine.h"



extern "C" {
void *hkg_device_alloc(uint64_t bytes)
{
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return p;
}
void hkg_device_free(void *p)
{
  if (p) (void)hipFree(p);
}
hkg_status hkg_memcpy(void *dst, const void *src, uint64_t bytes, int kind)
{
  hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  HKG_HIP_TRY(hipMemcpy(dst, src, bytes, k));
  return HKG_OK;
}
hkg_status hkg_device_sync(void)
{
  HKG_HIP_TRY(hipDeviceSynchronize());
  return HKG_OK;
}
}
