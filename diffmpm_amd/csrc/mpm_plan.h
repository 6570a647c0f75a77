// Internal definition of the opaque dmpm_plan (one particle set on one mesh).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/diffmpm_amd.h"

struct dmpm_plan {
  dmpm_plan_desc d{};
  int64_t n = 0;        // particles
  int dim = 2;
  int ncells = 0;
  int nnodes = 0;
  int device = 0;
  // counting-sort workspace
  int* key = nullptr;     // [n] cell of each particle, ncells = outside the mesh
  int* perm = nullptr;    // [n] sorted slot -> particle index
  int* counts = nullptr;  // [ncells+1]
  int* fill = nullptr;    // [ncells+1]
  int* cs = nullptr;      // [ncells+2] first slot of each cell (+ total)
  int* bsum = nullptr;    // scan block totals
  int* flags = nullptr;   // [0] sort dirty
  // velocity constraints, last value per (node, dir)
  int ncon = 0;
  int* con_node = nullptr;
  int* con_dir = nullptr;
  double* con_val = nullptr;
  // concentrated nodal forces grouped by (node, dir)
  int nnf = 0;
  int *nf_id = nullptr, *nf_dir = nullptr, *nf_off = nullptr, *nf_fn = nullptr;
  double* nf_val = nullptr;
  // particle tractions grouped by (particle, dir)
  int ntr = 0;
  bool has_tractions = false;
  int *tr_id = nullptr, *tr_dir = nullptr, *tr_off = nullptr, *tr_fn = nullptr;
  double* tr_val = nullptr;
  // time functions
  int nfn = 0;
  int *fn_kind = nullptr, *fn_off = nullptr, *fn_cnt = nullptr;
  double *fn_x = nullptr, *fn_f = nullptr;
  // workspace of the fused tile kernels and the adjoint (allocated lazily)
  void* extra[16] = {nullptr};
  size_t extra_bytes[16] = {0};

  void release_extra() {
    for (int i = 0; i < 16; ++i)
      if (extra[i]) {
        hipFree(extra[i]);
        extra[i] = nullptr;
        extra_bytes[i] = 0;
      }
  }
};

namespace dmpm {
int fail(int code, const std::string& msg);
int hip_check(hipError_t e, const char* what);
int launch_check(const char* what);
int enqueue_sort(dmpm_plan* P, hipStream_t s);
int ensure_sorted(dmpm_plan* P, const int32_t* eid, hipStream_t s);
// Returns extra workspace slot `slot` with at least `bytes` bytes.
void* plan_workspace(dmpm_plan* P, int slot, size_t bytes);
}  // namespace dmpm
