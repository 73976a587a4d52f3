// Internal definitions shared by the HIP translation units of liblfr.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <functional>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/lfr.h"

// An instantiated solver graph.  `key` holds every pointer and scalar the
// captured kernels were given, with library handles named by their uid;
// `deps` lists the uids of the handles whose device buffers the graph reads
// (plans, packed operators, spaces), so destroying one of them evicts it.
struct lfr_graph_entry {
  std::vector<uint64_t> key;
  std::vector<uint64_t> deps;
  hipGraphExec_t exec = nullptr;
};

// Process-wide identifier for library handles; never reused.
uint64_t lfr_next_uid();
// Destroy every cached graph of ctx that depends on the handle `uid`.
void lfr_graphs_evict(lfr_ctx* ctx, uint64_t uid);

struct lfr_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  std::string err;
  // grow-only device workspace for reductions and solver scalars
  void* work = nullptr;
  size_t work_bytes = 0;
  // pinned host staging for scalar read-backs (64 doubles)
  double* host_scalars = nullptr;
  int cu_count = 256;
  // events marking scalar read-backs (solver chunks in flight)
  hipEvent_t ev[2] = {nullptr, nullptr};
  // instantiated solver graphs, most recently used last
  std::vector<lfr_graph_entry> graphs;
  // grow-only device workspace of the generic iterative solvers (their
  // Krylov vectors, scalars and partial sums), apart from `work` so that an
  // operator or preconditioner applied inside a solve may use `work`
  void* kwork = nullptr;
  size_t kwork_bytes = 0;
  double solve_bytes = -1.0;        // algorithmic bytes of the last lfr_iterative_solve (-1: unknown)
  // autotuner: kept variant per (kernel family, shape) key, and its events
  std::map<std::string, int> tune;
  hipEvent_t tev[2] = {nullptr, nullptr};
};

struct lfr_space {
  lfr_ctx* ctx = nullptr;
  uint64_t uid = 0;                 // unique per created space (never reused)
  int ndf = 0, ncell = 0, nlayers = 0, undf = 0;
  int max_first = 0;                // largest map(1, cell) (1-based) over the cells
  int32_t* map0_d = nullptr;        // map(ndf, ncell) converted to 0-based, device
  std::vector<int32_t> map_h;       // 1-based host copy (plan construction)
  std::vector<int32_t> boundary_h;  // boundary_value(ndf, 2) for enforce_bc (empty: not set)
  mutable int vstack = -1;          // every cell's dof j = its dof 1 + j - 1 (one column; -1: not checked)
  int col_stride = -1;              // map(1, c) - 1 == c * col_stride for every cell (-1: not so)
};

// Queue enforce_bc over all cells of `sp` on `field` (sp must have its
// boundary values set).
int lfr_enforce_bc_queue(lfr_ctx* ctx, int prec, void* field, const lfr_space* sp);

// One class of output-dof columns with identical structure: every column has
// `height` dofs and its contributions come from `ngroup` cells, each
// contributing `gsize` local dofs (entries), listed in summation order.
struct lfr_plan_class {
  int height = 0;
  int ngroup = 0;
  int gsize = 0;
  int ncol = 0;
  int32_t* col_start_d = nullptr;   // (ncol) first dof of the column, 0-based
  int32_t* ent_d = nullptr;         // (ncol, ngroup*gsize, 2): cell (0-based), (i << 1) | shift
  std::vector<int32_t> starts_h;    // host copies, for compiling column records
  std::vector<int32_t> ent_h;
};

struct lfr_plan {
  lfr_ctx* ctx = nullptr;
  uint64_t uid = 0;                 // unique per created plan (never reused)
  const lfr_space* space = nullptr;
  int kind = 0;
  int ncolumns = 0;
  int64_t nbytes = 0;
  std::vector<lfr_plan_class> classes;
  // compiled column records per input space (one device array per class,
  // nullptr where the class shape has no record kernel)
  // (keyed by the input space's uid, which is never reused)
  mutable std::vector<std::pair<uint64_t, std::vector<int32_t*>>> records;
  // records of the unified memory-order gather, per input space
  mutable std::vector<std::pair<uint64_t, int32_t*>> urecords;
};

// Column records of `plan` class `ci` for input space `in`, built on first
// use; nullptr when the shape has no record kernel.  Returns the int stride.
const int32_t* lfr_plan_records(lfr_ctx* ctx, const lfr_plan* plan, size_t ci, const lfr_space* in,
                                int* stride);
bool lfr_record_shape(int ndf_in, int ngroup, int gsize);
bool lfr_records_enabled();
bool lfr_gather_vec_nt(bool packed);
// The plan qualifies for the unified memory-order gather (every column has
// two entries and nl or nl+1 levels; ndf_in 1 or 6; nl >= lpt).
bool lfr_plan_unified(const lfr_plan* plan, int ndf_in, int lpt);
int lfr_plan_ncolumns_total(const lfr_plan* plan);
int lfr_unified_lpt(int prec);
bool lfr_gather_stage();
// Unified records for input space `in` (built on first use), in column
// memory order; nullptr on allocation failure (error set on ctx).
const int32_t* lfr_plan_urecords(lfr_ctx* ctx, const lfr_plan* plan, const lfr_space* in);
// (cell, code) of entries 0 and 1 of every column in memory order (host).
void lfr_plan_uentries(const lfr_plan* plan, std::vector<int32_t>& ent);
// Blocks of the unified launch for levels-per-thread lpt.
unsigned lfr_unified_blocks(const lfr_plan* plan, int lpt);
bool lfr_gather_nt(int prec);
// Levels per thread of the record gather for plan class `c` (1: one level per
// thread; 2 or 4: the vectorised kernel, used for classes of exactly nl
// levels with a record shape); blocks of 256 threads that class launches.
int lfr_class_lpt(const lfr_plan_class& c, int nl, int prec, int ndf_in, bool packed = false);
unsigned lfr_class_blocks(const lfr_plan_class& c, int lpt);

// Packed operator: the matrix rows each output column of `plan` reads,
// stored per column in the plan's order, entry-major, then input dof, then
// level (class heights), zero where an entry has no cell layer.
struct lfr_lma_packed {
  lfr_ctx* ctx = nullptr;
  uint64_t uid = 0;                 // unique per packed operator (never reused)
  int prec = 0;
  int kind = 0;
  const lfr_plan* plan = nullptr;
  const lfr_space* in = nullptr;
  int ndf_out = 0;
  void* data_d = nullptr;
  std::vector<int64_t> off;  // element offset of each class's block
  std::vector<int64_t> pcs;  // elements per column of each class
  int64_t nbytes = 0;
  // unified layout: every column in memory order, 2 entries x ndf_in rows
  // of nl+1 levels (the unified gather's packed form); off/pcs have one slot
  bool unified = false;
  // unified gather variant kept by the autotuner at pack time (-1: the
  // global default): output staged through LDS, non-temporal matrix loads
  int stage = -1, nt = -1;
};

int lfr_set_error(lfr_ctx* ctx, int code, const char* fmt, ...);

// Autotuner (lfr_tune.hip): the stored choice for `key`, else (tuning on,
// stream not capturing) the fastest of ncand result-identical variants,
// timed by calling run(c) on ctx->stream, which is stored; else dflt.
int lfr_tune_pick(lfr_ctx* ctx, const std::string& key, int ncand, int dflt, const std::function<int(int)>& run);
bool lfr_tune_enabled();
int lfr_tune_load_file(lfr_ctx* ctx, const char* path);
int lfr_tune_save_file(const lfr_ctx* ctx, const char* path);
// Unified-gather variant flags of a packed operator (its tuned choice or the
// global default).
bool lfr_pk_stage(const lfr_lma_packed* pk);
bool lfr_pk_nt(const lfr_lma_packed* pk);

// Validate a square solver operator (kind-0 plan of its space, packed copy
// of the same plan) and build its column records outside any graph capture.
int lfr_check_op(lfr_ctx* ctx, const lfr_lma_op* A, const char* who);
// Queue y = A x for a checked square operator (enforce_bc after the apply
// when A->apply_bc), as mass_matrix_operator_type%apply.
int lfr_lma_apply_queue(lfr_ctx* ctx, const lfr_lma_op* A, const void* x, void* y);
// The same with the partial sums of <x, A x> (one per apply workgroup, in a
// fixed order) written to part[0 .. *npart); part must hold
// lfr_lma_apply_partials_max(A->plan) doubles.
int lfr_lma_apply_dot_queue(lfr_ctx* ctx, const lfr_lma_op* A, const void* x, void* y, double* part,
                            int* npart);
int lfr_lma_apply_partials_max(const lfr_plan* plan);

// Output and input field sizes (undf) of a multi-field operator handle.
extern "C" int lfr_multi_space_undf(const lfr_multi* h, int* undf_to, int* undf_from);

#define LFR_HIP(ctx, call)                                                        \
  do {                                                                            \
    hipError_t _e = (call);                                                       \
    if (_e != hipSuccess)                                                         \
      return lfr_set_error((ctx), LFR_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, \
                           #call, hipGetErrorString(_e));                         \
  } while (0)

#define LFR_CHECK_LAUNCH(ctx) LFR_HIP(ctx, hipGetLastError())

#define LFR_REQUIRE(ctx, cond, ...)                       \
  do {                                                    \
    if (!(cond)) return lfr_set_error((ctx), LFR_ERR_ARG, __VA_ARGS__); \
  } while (0)

// Resolve the device address of the allocation-free entry points' scratch
// for ctx->device (lfr_init calls it with that device current).
int lfr_noalloc_setup(lfr_ctx* ctx);

// Grow the context workspace to at least `bytes`.
int lfr_work_reserve(lfr_ctx* ctx, size_t bytes);

// Blocks for a grid-stride launch over n items with `threads` per block.
static inline unsigned lfr_blocks(int64_t n, int threads, unsigned cap = 1u << 30) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > (int64_t)cap) b = cap;
  return (unsigned)b;
}

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs, each
// with a private L2.  Remapping the hardware block id so that each XCD gets a
// contiguous range of logical blocks keeps neighbouring columns (which share
// gathered operand dofs) in one L2.  nb_padded must be a multiple of 8.
__device__ __forceinline__ unsigned lfr_xcd_block(unsigned hw_block, unsigned nb_padded) {
  const unsigned per = nb_padded >> 3;
  return (hw_block & 7u) * per + (hw_block >> 3);
}
