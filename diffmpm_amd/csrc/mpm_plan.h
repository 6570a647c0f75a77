// Internal definition of the opaque dmpm_plan (one particle set on one mesh).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/diffmpm_amd.h"

// Cells per side of the square tiles of the one-launch step.
constexpr int kStepTile = 16;

// One simulation of a plan (device table; 64 bytes).
struct SimRec {
  int nx, ny;                         // cells per axis
  int cell_off, node_off, part_off;   // first global cell, node, particle
  int n;                              // particles
  int tile_off, tiles_x;              // one-launch step: first global tile, tile columns
  double hx, hy;
  double de[3];                       // material coefficients
  double pad2;
};

struct dmpm_plan {
  dmpm_plan_desc d{};
  int64_t n = 0;        // particles
  int dim = 2;
  int ncells = 0;   // all simulations (a batch plan holds several)
  int nnodes = 0;
  int nsims = 1;
  SimRec* sims = nullptr;          // [nsims] device table
  std::vector<SimRec> sims_h;      // host copy
  int ntiles = 0;                  // tiles of the one-launch step, all simulations
  int64_t* clock = nullptr;        // device step counter (dmpm_run's DMPM_RUN_CLOCK)
  int* tile_sim = nullptr;         // [ntiles] simulation of each tile (batch plans)
  int device = 0;
  // counting-sort workspace
  int* key = nullptr;     // [n] cell of each particle, ncells = outside the mesh
  int* perm = nullptr;    // [n] sorted slot -> particle index
  int* counts = nullptr;  // [ncells+1]
  int* fill = nullptr;    // [ncells+1]
  int* cs = nullptr;      // [ncells+2] first slot of each cell (+ total)
  int* bsum = nullptr;    // scan block totals
  // flags: [0] full re-sort needed, [1] perm is the identity, [2] and [7]
  // zero (a set [2] forces a full re-sort), [3] big-cell count, [4] current
  // buffer of perm/cs (0 or 1), [5] cell changes recorded since the last sort,
  // [6] a recorded change overflowed its list bound (full re-sort needed),
  // [8] finished workgroups of the last full-sort stage, [10] the per-phase
  // sort rebuilt buffer 0 since the last fused-path sort
  int* flags = nullptr;
  int* zflags = nullptr;  // [kFlagWords] zeros (flags of a caller-held sort order)
  int* big = nullptr;     // [ncells] cells queued for the big-cell sort
  int* perm2 = nullptr;   // second buffer of perm (incremental re-sort)
  int* cs2 = nullptr;     // second buffer of cs
  // recorded cell changes, kMovRec ints each: particle, old key, new element
  // id (of the particle's own mesh), next leaver of the old cell, next arrival
  // of the new cell, new key
  int movcap = 0;
  int* mov = nullptr;
  int* mrec = nullptr;    // [n] 1 + record index of a recorded particle, else 0
  int* dL = nullptr;      // [ncells+1] leavers per cell (zero between sorts)
  int* dA = nullptr;      // [ncells+1] arrivals per cell (zero between sorts)
  int* headL = nullptr;   // [ncells+1] leaver list heads (-1 between sorts)
  int* headA = nullptr;   // [ncells+1] arrival list heads (-1 between sorts)
  int nchunks = 0;        // cell chunks of 1 << chunk_shift cells (scan granularity)
  int chunk_shift = 10;
  int* chunkD = nullptr;  // [nchunks] net arrivals per chunk (zero between sorts)
  int* chunkC = nullptr;  // [nchunks] particles per chunk in a full sort (zero between)
  int nsub = 0;           // sub-chunks of kSubCells cells
  int* subT = nullptr;    // [nsub] sub-chunk has recorded cell changes (incremental re-sort)
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
  // dense per-node views of the constraint and nodal-force tables for the
  // tile kernels: cidx[node] -> record (cmask bit d: dof d fixed to cval[2r+d]),
  // nfd[d*nnodes + node] -> nodal-force group or -1
  int* cidx = nullptr;
  int* cmask = nullptr;
  double* cval = nullptr;
  int* nfd = nullptr;
  // time functions
  int nfn = 0;
  int *fn_kind = nullptr, *fn_off = nullptr, *fn_cnt = nullptr;
  double *fn_x = nullptr, *fn_f = nullptr;
  // per-kernel-class timing (dmpm_timing_enable): event pairs recorded around
  // launches of the enabled kernel classes on their streams, a ring of event
  // pairs per class
  struct ProfClass {
    bool on = false;
    int n = 0;
    double ms = 0;
    int pending[4] = {0, 0, 0, 0};
    hipEvent_t ev[8] = {};
  };
  ProfClass prof[16];
  int prof_single = -1;  // the class dmpm_timing_read reports (-2: all enabled)
  int cell_block = 128;  // cells per workgroup of the P2G cell pass (64/128/256)
  int usf_split = 1;     // USF: stress update streamed (1) or in the force cell pass (0)
  int step_kernel = 1;   // single-simulation fused step: one launch (1) or one per phase (0)
  unsigned char step_aux[256];  // host copy of the one-launch step's device-side tables
  unsigned char step_aux_ext[256];  // the same, as uploaded into caller-supplied memory
  int step_aux_bytes = 0;       // bytes of step_aux uploaded (0: none yet)
  // workspace of the fused tile kernels and the adjoint (allocated lazily)
  void* extra[16] = {nullptr};
  size_t extra_bytes[16] = {0};

  void release_extra() {
    for (int i = 0; i < 16; ++i)
      if (extra[i]) {
        (void)hipFree(extra[i]);  // release: nothing to do on failure
        extra[i] = nullptr;
        extra_bytes[i] = 0;
      }
  }
};

#define DMPM_TRY(x)                   \
  do {                                \
    int rc__ = (x);                   \
    if (rc__ != DMPM_OK) return rc__; \
  } while (0)
#define DMPM_TRY_H(x) DMPM_TRY(::dmpm::hip_check((x), #x))

namespace dmpm {
constexpr int kMovRec = 6;     // ints per recorded cell change
constexpr int kListMax = 64;   // recorded leavers/arrivals per cell handled incrementally
constexpr int kMinChunkShift = 10;     // at least 1024 cells (4 per thread) per scan chunk
constexpr int kSubShift = 6;           // sub-chunks of 64 cells (one wave) in the incremental re-sort
constexpr int kMaxChunks = 8192;       // LDS histogram bound of the full sort
constexpr int kFlagWords = 16;

// The current sort buffers (selected on the device by flags[4]).
struct SortRef {
  const int* perm0;
  const int* perm1;
  const int* cs0;
  const int* cs1;
  const int* flags;
  __device__ __forceinline__ const int* perm() const { return flags[4] ? perm1 : perm0; }
  __device__ __forceinline__ const int* cs() const { return flags[4] ? cs1 : cs0; }
};
inline SortRef sort_ref(const dmpm_plan* P) {
  SortRef r;
  r.perm0 = P->perm;
  r.perm1 = P->perm2;
  r.cs0 = P->cs;
  r.cs1 = P->cs2;
  r.flags = P->flags;
  return r;
}

int fail(int code, const std::string& msg);
int hip_check(hipError_t e, const char* what);
int launch_check(const char* what);
int enqueue_sort(dmpm_plan* P, hipStream_t s);
int ensure_sorted(dmpm_plan* P, const int32_t* eid, hipStream_t s);
// Returns extra workspace slot `slot` with at least `bytes` bytes (allocated
// on first use, kept for the plan's lifetime).
void* plan_workspace(dmpm_plan* P, int slot, size_t bytes);
// Node-tile step (mpm_tile.hip): 2D, no particle tractions or particle f_ext.
// clock: the steps' indices are *clock + 0 .. nsteps-1, read on the device,
// and the run adds nsteps to *clock at its end (null: step0 .. on the host).
int tile_run(dmpm_plan* P, void* pbuf, int32_t* eid, void* nbuf, int scheme, int vu,
             const double* grav, double dt, int64_t step0, int64_t nsteps, int temps,
             hipStream_t s, int64_t* clock);
// Allocates the workspaces of tile_run and uploads its tables (no other work).
int tile_prepare(dmpm_plan* P, int scheme, hipStream_t s);
bool tile_eligible(const dmpm_plan* P, int pfext);
// Copies the sort by cell of the plan's last step to perm_out [n] and cs_out
// [ncells + 2] (mpm_adjoint.hip).
int export_sort(dmpm_plan* P, int32_t* perm_out, int32_t* cs_out, hipStream_t s);
// One fused 2D step from block pin to block pout (mpm_tile.hip).
int tile_between(dmpm_plan* P, const void* pin, void* pout, int32_t* eid, void* nbuf,
                 int scheme, int vu, const double* grav, double dt, int64_t step, int temps,
                 int chained, const void* sin, void* sout, int64_t sld, hipStream_t s);
// tile_between of a differentiable rollout's saved step (dmpm_step_between_saved).
int tile_between_saved(dmpm_plan* P, const void* pin, void* pout, int32_t* eid,
                       const void* nprev, void* nout, int scheme, int vu, const double* grav,
                       double dt, int64_t step, int chained, const void* sin, void* sout,
                       int64_t sld, void* vkin_out, int32_t* perm_out, int32_t* cs_out,
                       hipStream_t s);
// dmpm_step_between on workspaces in the library's static device array
// (mpm_tile.hip); noalloc_scratch_setup takes that array's address on a
// device (once per device, with the device current).
int tile_between_noalloc(dmpm_plan* P, const void* pin, void* pout, int32_t* eid, void* nbuf,
                         int scheme, int vu, const double* grav, double dt, int64_t step,
                         int temps, int chained, const void* sin, void* sout, int64_t sld,
                         hipStream_t s);
int noalloc_scratch_setup(int device);
// One per-phase step (mpm_forward.hip) that also copies the kinematic nodal
// velocity (dim rows of nnodes) to vkin.
int phase_step_keep_vkin(dmpm_plan* P, void* pbuf, int32_t* eid, void* nbuf, int scheme, int vu,
                         const double* grav, double dt, int64_t step, void* vkin, hipStream_t s);
// The adjoint of one Linear1D step (mpm_adjoint_1d.hip); arguments as
// dmpm_step_adjoint.
int adjoint_1d(dmpm_plan* P, const void* pin, void* pwork, int32_t* eid, void* nwork, void* cot,
               void* lebar, int scheme, int vu, const double* grav, double dt, int64_t step,
               hipStream_t s);
// Kernel classes for dmpm_timing_enable.
enum { KC_LOCATE = 0, KC_SORT = 1, KC_P2G_KIN = 2, KC_STRESS = 3, KC_P2G_FORCE = 4,
       KC_GRID = 5, KC_G2P = 6, KC_SMALL = 7, KC_TILE_KIN = 8, KC_TILE_FORCE = 9,
       KC_TILE_G2P = 10, KC_ADJ = 11, KC_BATCH = 12, KC_STEP = 13 };
constexpr int kNumKC = 14;
// A ring of kProfRing event pairs per class: before a pair is reused its
// elapsed time is read (host waits for the launch of that class kProfRing
// launches back), so at most kProfRing timed launches per class are
// outstanding.
constexpr int kProfRing = 4;
inline void prof_drain_slot(dmpm_plan::ProfClass& c, int slot) {
  if (!c.pending[slot]) return;
  (void)hipEventSynchronize(c.ev[2 * slot + 1]);
  float ms = 0;
  if (hipEventElapsedTime(&ms, c.ev[2 * slot], c.ev[2 * slot + 1]) == hipSuccess) c.ms += ms;
  c.pending[slot] = 0;
}
inline void prof_begin(dmpm_plan* P, int cls, hipStream_t s) {
  dmpm_plan::ProfClass& c = P->prof[cls];
  if (!c.on) return;
  const int slot = c.n % kProfRing;
  prof_drain_slot(c, slot);
  (void)hipEventRecord(c.ev[2 * slot], s);
}
inline void prof_end(dmpm_plan* P, int cls, hipStream_t s) {
  dmpm_plan::ProfClass& c = P->prof[cls];
  if (!c.on) return;
  const int slot = c.n % kProfRing;
  (void)hipEventRecord(c.ev[2 * slot + 1], s);
  c.pending[slot] = 1;
  c.n++;
}
// The event pair of the next timed launch of class cls (false when the
// class is not timed): the caller records them itself (for example as the
// start and stop events of one kernel dispatch) and then calls prof_commit.
inline bool prof_slot(dmpm_plan* P, int cls, hipEvent_t* e0, hipEvent_t* e1) {
  dmpm_plan::ProfClass& c = P->prof[cls];
  if (!c.on) return false;
  const int slot = c.n % kProfRing;
  prof_drain_slot(c, slot);
  *e0 = c.ev[2 * slot];
  *e1 = c.ev[2 * slot + 1];
  return true;
}
inline void prof_commit(dmpm_plan* P, int cls) {
  dmpm_plan::ProfClass& c = P->prof[cls];
  c.pending[c.n % kProfRing] = 1;
  c.n++;
}
struct ProfScope {
  dmpm_plan* P;
  int cls;
  hipStream_t s;
  ProfScope(dmpm_plan* p, int c, hipStream_t st) : P(p), cls(c), s(st) { prof_begin(P, cls, s); }
  ~ProfScope() { prof_end(P, cls, s); }
};
}  // namespace dmpm
using ProfScopeT = dmpm::ProfScope;
