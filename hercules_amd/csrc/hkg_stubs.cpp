// Entry points whose device implementations are not in this build yet.
#include <string.h>

#include "hkg_engine.h"



extern "C" {
void hkg_default_server_opts(hkg_server_opts *so)
{
  if (so) memset(so, 0, sizeof(*so));
}
hkg_status hkg_server_start(hkg_index *, const hkg_server_opts *, hkg_server **) { return HKG_EINVAL; }
hkg_status hkg_submit(hkg_server *, const float *, uint32_t, uint64_t *) { return HKG_EINVAL; }
hkg_status hkg_poll(hkg_server *, uint64_t, uint64_t *, float *, double *) { return HKG_EINVAL; }
hkg_status hkg_wait(hkg_server *, uint64_t, uint64_t *, float *, double *) { return HKG_EINVAL; }
void hkg_server_stop(hkg_server *) {}

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
