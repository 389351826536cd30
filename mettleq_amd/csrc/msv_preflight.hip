// Statevector memory preflight (the decision of the reference's
// statevector_preflight, execution.py:270-374, re-sized for one MI355X and
// this engine's in-place kernels) and the device limits it checks against.
//
// Two lower bounds exist: out of place (two states, the reference's model,
// execution.py:195-223) for circuits with XOR permutations, whose gather
// writes a second buffer, and in place (one state plus the library's fixed
// scratch) for everything else.
#include <hip/hip_runtime.h>
#include <hsa/hsa.h>
#include <hsa/hsa_ext_amd.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/mettleq_sv.h"
#include "msv_internal.h"

namespace {

constexpr int kQubitLimit = 40;                   // 64-bit indices; largest state the ABI creates
constexpr uint64_t kFixedScratch = 8ull << 20;    // staging ring + reduction scratch + tables

struct PoolQuery {
  uint32_t bdf = 0;
  uint32_t domain = 0;
  bool found = false;
  uint64_t max_alloc = 0;
};

hsa_status_t pool_cb(hsa_amd_memory_pool_t pool, void *data) {
  auto *q = static_cast<PoolQuery *>(data);
  hsa_amd_segment_t seg;
  if (hsa_amd_memory_pool_get_info(pool, HSA_AMD_MEMORY_POOL_INFO_SEGMENT, &seg) !=
          HSA_STATUS_SUCCESS ||
      seg != HSA_AMD_SEGMENT_GLOBAL)
    return HSA_STATUS_SUCCESS;
  bool alloc_ok = false;
  if (hsa_amd_memory_pool_get_info(pool, HSA_AMD_MEMORY_POOL_INFO_RUNTIME_ALLOC_ALLOWED,
                                   &alloc_ok) != HSA_STATUS_SUCCESS ||
      !alloc_ok)
    return HSA_STATUS_SUCCESS;
  uint32_t flags = 0;
  if (hsa_amd_memory_pool_get_info(pool, HSA_AMD_MEMORY_POOL_INFO_GLOBAL_FLAGS, &flags) !=
          HSA_STATUS_SUCCESS ||
      !(flags & HSA_AMD_MEMORY_POOL_GLOBAL_FLAG_COARSE_GRAINED))
    return HSA_STATUS_SUCCESS;
  size_t mx = 0;
  if (hsa_amd_memory_pool_get_info(pool, HSA_AMD_MEMORY_POOL_INFO_ALLOC_MAX_SIZE, &mx) ==
      HSA_STATUS_SUCCESS)
    q->max_alloc = std::max<uint64_t>(q->max_alloc, mx);
  return HSA_STATUS_SUCCESS;
}

hsa_status_t agent_cb(hsa_agent_t agent, void *data) {
  auto *q = static_cast<PoolQuery *>(data);
  hsa_device_type_t type;
  if (hsa_agent_get_info(agent, HSA_AGENT_INFO_DEVICE, &type) != HSA_STATUS_SUCCESS ||
      type != HSA_DEVICE_TYPE_GPU)
    return HSA_STATUS_SUCCESS;
  uint32_t bdf = 0, dom = 0;
  if (hsa_agent_get_info(agent, static_cast<hsa_agent_info_t>(HSA_AMD_AGENT_INFO_BDFID), &bdf) !=
          HSA_STATUS_SUCCESS ||
      hsa_agent_get_info(agent, static_cast<hsa_agent_info_t>(HSA_AMD_AGENT_INFO_DOMAIN), &dom) !=
          HSA_STATUS_SUCCESS)
    return HSA_STATUS_SUCCESS;
  if (bdf != q->bdf || dom != q->domain) return HSA_STATUS_SUCCESS;
  q->found = true;
  (void)hsa_amd_agent_iterate_memory_pools(agent, pool_cb, data);
  return HSA_STATUS_INFO_BREAK;
}

std::mutex g_alloc_mu;
uint64_t g_alloc_max[64];
bool g_alloc_done[64];

std::string gib(uint64_t v, bool known) {
  if (!known) return "unknown";
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.2f GiB", static_cast<double>(v) / (1024.0 * 1024.0 * 1024.0));
  return buf;
}

}  // namespace

namespace msv {

// Largest single allocation the runtime allows on `device`: ALLOC_MAX_SIZE of
// the HSA agent's coarse-grained global memory pool (the pool hipMalloc
// draws from), matched to the HIP device by PCI domain and bus/device/
// function. 0 when the runtime does not report it.
uint64_t device_max_alloc(int device) {
  if (device < 0 || device >= 64) return 0;
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  if (g_alloc_done[device]) return g_alloc_max[device];
  g_alloc_done[device] = true;
  char bus_id[64] = {0};
  if (hipDeviceGetPCIBusId(bus_id, sizeof(bus_id), device) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  unsigned dom = 0, bus = 0, dev = 0, fn = 0;
  if (std::sscanf(bus_id, "%x:%x:%x.%x", &dom, &bus, &dev, &fn) != 4) return 0;
  PoolQuery q;
  q.bdf = (bus << 8) | (dev << 3) | fn;
  q.domain = dom;
  if (hsa_init() != HSA_STATUS_SUCCESS) return 0;
  (void)hsa_iterate_agents(agent_cb, &q);
  (void)hsa_shut_down();
  g_alloc_max[device] = q.found ? q.max_alloc : 0;
  return g_alloc_max[device];
}

}  // namespace msv

extern "C" {

int msv_preflight(int n_qubits, int dtype, int mode, uint64_t reserved_bytes, int allow_unsafe,
                  int device, const msv_preflight_limits *limits, msv_preflight_report *out) {
  if (!out) return msv::set_error(MSV_E_INVALID, "preflight: null report");
  std::memset(out, 0, sizeof(*out));
  if (n_qubits < 0) return msv::set_error(MSV_E_INVALID, "n_qubits must be non-negative");
  if (dtype != MSV_C64 && dtype != MSV_C128)
    return msv::set_error(MSV_E_INVALID, "preflight: unknown dtype");
  if (mode != MSV_PF_OUT_OF_PLACE && mode != MSV_PF_IN_PLACE)
    return msv::set_error(MSV_E_INVALID, "preflight: mode must be out of place or in place");

  const uint64_t bpa = dtype == MSV_C64 ? 8 : 16;
  // saturate above 2^63 bytes (such widths are refused by the qubit limit anyway)
  const uint64_t state = n_qubits >= 59 ? ~0ull : (1ull << n_qubits) * bpa;
  const uint64_t lower = mode == MSV_PF_OUT_OF_PLACE
                             ? (state > (~0ull >> 1) ? ~0ull : 2 * state)
                             : (state > ~0ull - kFixedScratch ? ~0ull : state + kFixedScratch);

  uint64_t max_buffer = 0, working = 0;
  if (limits) {
    max_buffer = limits->max_buffer_bytes;
    working = limits->working_set_bytes;
  } else {
    // live limits of `device`: largest allocation and free memory now (which
    // already excludes every state this process holds)
    if (hipSetDevice(device) == hipSuccess) {
      size_t fr = 0, tot = 0;
      if (hipMemGetInfo(&fr, &tot) == hipSuccess) working = fr;
      max_buffer = msv::device_max_alloc(device);
    }
    (void)hipGetLastError();
  }
  const bool have_buffer = max_buffer > 0;
  const bool have_working = working > 0;
  const uint64_t avail = working > reserved_bytes ? working - reserved_bytes : 0;

  out->state_bytes = state;
  out->lower_bound_bytes = lower;
  out->max_buffer_bytes = have_buffer ? max_buffer : 0;
  out->working_set_bytes = have_working ? avail : 0;
  out->buffer_check = have_buffer ? (state <= max_buffer ? 1 : 0) : -1;
  out->working_check = have_working ? (lower <= avail ? 1 : 0) : -1;
  int failures = 0;
  if (out->buffer_check == 0) failures |= MSV_PF_FAIL_MAX_BUFFER;
  if (out->working_check == 0) failures |= MSV_PF_FAIL_WORKING_SET;
  if (n_qubits > kQubitLimit) failures |= MSV_PF_FAIL_QUBITS;
  out->failures = failures;
  const bool refused = failures != 0;
  const bool overridden = refused && allow_unsafe != 0 && !(failures & MSV_PF_FAIL_QUBITS);
  out->overridden = overridden ? 1 : 0;
  out->allowed = (!refused || overridden) ? 1 : 0;
  if (overridden)
    out->decision = MSV_PF_ALLOWED_OVERRIDE;
  else if (refused)
    out->decision = MSV_PF_REFUSED;
  else if (have_buffer && have_working)
    out->decision = MSV_PF_ALLOWED;
  else
    out->decision = MSV_PF_ALLOWED_UNVERIFIED;
  if (out->allowed) return MSV_OK;

  const char *ty = dtype == MSV_C64 ? "complex64" : "complex128";
  const char *bound = mode == MSV_PF_OUT_OF_PLACE ? "out-of-place" : "in-place";
  std::string msg = "Refusing " + std::to_string(n_qubits) + "-qubit " + ty +
                    " statevector before allocation: one state is " + gib(state, true) +
                    " and the " + bound + " lower bound is " + gib(lower, true) +
                    "; device max buffer is " + gib(max_buffer, have_buffer) +
                    " and free working set is " + gib(avail, have_working) +
                    ". Reduce qubits, use complex64, or set METTLEQ_ALLOW_UNSAFE_STATEVECTOR=1 "
                    "to deliberately bypass this preflight.";
  if (failures & MSV_PF_FAIL_QUBITS)
    msg += " (" + std::to_string(n_qubits) + " qubits exceed the engine limit of " +
           std::to_string(kQubitLimit) + ")";
  return msv::set_error(MSV_E_REFUSED, msg);
}

}  // extern "C"
