// ctx.cpp -- context lifetime, errors and version of the nlmech_amd library.

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <new>
#include <string>

#include "nlm_internal.h"

namespace {
std::string g_create_error;
}

extern "C" int nlm_ctx_create(nlm_ctx** out, int device) {
  if (!out) return NLM_ERR_INVALID;
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) {
    g_create_error = "nlm_ctx_create: no HIP device available";
    return NLM_ERR_HIP;
  }
  if (device < 0 || device >= ndev) {
    g_create_error = "nlm_ctx_create: device index out of range";
    return NLM_ERR_INVALID;
  }
  if ((e = hipSetDevice(device)) != hipSuccess) {
    g_create_error = std::string("hipSetDevice: ") + hipGetErrorString(e);
    return NLM_ERR_HIP;
  }
  nlm_ctx* c = new (std::nothrow) nlm_ctx;
  if (!c) return NLM_ERR_MEMORY;
  c->device = device;
  {
    std::string msg;
    int rc = nlm::scratch_setup(device, &c->scratch, msg);
    if (rc != NLM_OK) {
      g_create_error = "nlm_ctx_create: " + msg;
      delete c;
      return rc;
    }
  }
  // default-priority stream
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) !=
      hipSuccess) {
    g_create_error = std::string("hipStreamCreate: ") + hipGetErrorString(e);
    delete c;
    return NLM_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess)
    c->arch = prop.gcnArchName;
  // NLM_QS_GRAPH=0 starts contexts with the BiCGSTAB batches launched kernel
  // by kernel (nlm_qs_set_graph); profilers that do not follow captured
  // graphs need it
  if (const char* g = std::getenv("NLM_QS_GRAPH"))
    c->qs.use_graph = std::strcmp(g, "0") != 0;
  *out = c;
  return NLM_OK;
}

extern "C" void nlm_ctx_destroy(nlm_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->fd.ev0) (void)hipEventDestroy(c->fd.ev0);
  if (c->fd.ev1) (void)hipEventDestroy(c->fd.ev1);
  if (c->fd.graph) (void)hipGraphExecDestroy(c->fd.graph);
  if (c->qs.bicg_graph) (void)hipGraphExecDestroy(c->qs.bicg_graph);
  if (c->qs.e0) (void)hipEventDestroy(c->qs.e0);
  if (c->qs.e1) (void)hipEventDestroy(c->qs.e1);
  if (c->qs.hctl) (void)hipHostFree(c->qs.hctl);
  hipStream_t st = c->stream;
  delete c;  // frees every device buffer
  if (st) (void)hipStreamDestroy(st);
}

extern "C" const char* nlm_last_error(const nlm_ctx* c) {
  if (!c) return g_create_error.c_str();
  return c->err.c_str();
}

extern "C" const char* nlm_version(void) { return "nlmech_amd 0.1.0 gfx950"; }

extern "C" int nlm_device_arch(const nlm_ctx* c, char* buf, int len) {
  if (!c || !buf || len <= 0) return NLM_ERR_INVALID;
  std::strncpy(buf, c->arch.c_str(), size_t(len - 1));
  buf[len - 1] = '\0';
  return NLM_OK;
}
