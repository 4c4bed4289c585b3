// plan.hpp -- host-side setup of the projection path: everything that is
// built once per design and uploaded to HBM (SolverParams / Solver::from_design,
// crates/within/src/solver.rs:98-132).
//
// The plan flattens the reference's per-subdomain objects into concatenated
// arrays so that one kernel launch serves every subdomain:
//   * observation layouts for the deterministic implicit operator
//     (level-sorted CSR per factor; domain.rs:303-432 semantics),
//   * cross-tabs / components / partition weights (factor_pairs.rs,
//     gramian/cross_tab.rs),
//   * Schur complements (schur_complement.rs) and the reduced-system factor
//     as rounds of independent eliminations plus a dense anchored core,
//   * the explicit Gramian CSR (gramian/explicit.rs).
#pragma once
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <mutex>
#include <vector>

namespace wa {

using u32 = uint32_t;
using i64 = int64_t;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

struct LocalConfig {
  uint64_t ac_seed = 0;
  int ac_split = 1;
  bool approx_schur = true;
  uint64_t schur_seed = 0;
  int schur_split = 1;
  int dense_threshold = 24;
  i64 dense_core = 16384;
  bool device = false;  // factor dense cores of at least kDeviceCore on the GPU
};
constexpr i64 kDeviceCore = 2048;

struct Design {
  i64 n_obs = 0;
  int Q = 0;
  std::vector<std::vector<u32>> codes;  // codes[q][i]
  std::vector<double> weights;          // empty = unit
  std::vector<i64> levels, offset;
  i64 n_dofs = 0;
  double w(i64 i) const { return weights.empty() ? 1.0 : weights[i]; }
};
Design make_design(const u32* codes, i64 n_obs, int Q, i64 ld, const double* weights);

// Work decomposition of a segmented reduction over rows [r0, r1) of a CSR:
// rows of at most `short_max` elements need no item (each is summed by one lane
// group); a longer row is one work item [row0, row1 = row0 + 1) (written
// directly) or, beyond `chunk` elements, several chunks [e0, e1) written to
// partial slots `slot` that a fix-up sums in slot order.
struct Item { u32 row0, row1, e0, e1; int32_t slot; int32_t pad; };
// A fix-up of a row with more than 2 * kFixChunk slots is done in two stages:
// blocks of kFixChunk slots are first summed into their own partial slots
// starting at chunk0, and the fix-up then sums those (kNoChunks: one stage).
struct Fixup { u32 row, slot0, slot1, chunk0; };
constexpr u32 kFixChunk = 256;
constexpr u32 kNoChunks = 0xFFFFFFFFu;
// one row range with its items [i0, i1) and fix-ups [f0, f1); max_chunks:
// the most first-stage blocks of any of its fix-ups (0: all one stage)
struct Segment { u32 r0, r1; i64 i0, i1, f0, f1; u32 max_chunks = 0; };

// Workgroups are dispatched to the MI355X's 8 XCDs (each with its own L2)
// round-robin, and the engine runs 4 waves (items) per workgroup.  Chunks of
// different split rows that cover the same stretch of columns are placed in
// workgroups with the same index mod 8, so they share one L2.
constexpr u32 kXcds = 8, kWavesPerGroup = 4;
struct Schedule {
  std::vector<Item> items;
  std::vector<Fixup> fixups;
  std::vector<Segment> segs;
  i64 n_slots = 0;
  u32 short_max = 0;
};
// rows r with skip[r] set get no item (another kernel computes them; they
// must be longer than short_max)
Schedule make_schedule(const std::vector<u32>& row_ptr, u32 chunk, u32 short_max,
                       const std::vector<char>* skip = nullptr);
// one more segment, rows [r0, r1) (slots continue from s.n_slots)
void append_schedule(Schedule& s, const std::vector<u32>& row_ptr, u32 r0, u32 r1, u32 chunk,
                     const std::vector<char>* skip = nullptr);

// Level-sorted CSR of one factor: rows = levels of q, elements = observations
// sorted by (level, code of the largest other factor, observation index).
struct FactorPass {
  int q = 0;
  std::vector<u32> row_ptr;   // levels[q] + 1
  std::vector<u32> obs;       // element -> observation
  std::vector<u32> other;     // element-major, (Q-1) global DOF indices each
  std::vector<double> w;      // element weights (empty = unit)
  std::vector<double> diag;   // sum of weights per level
};

struct Csr {
  i64 nrows = 0, ncols = 0;
  std::vector<u32> indptr, indices;
  std::vector<double> data;
  Csr transpose() const;
};

struct CrossTab {
  Csr c, ct;
  std::vector<double> diag_q, diag_r;
  i64 n_q() const { return c.nrows; }
  i64 n_r() const { return c.ncols; }
};

// A dense core's anchored inverse kept in device memory (FP64, n x n
// row-major) where it was computed; host copies are made only on request
// (serialisation, exports).  Shared by the objects of one plan.  Its last
// kernels may still run on a stream of their own when it is handed over, so
// every reader of p calls ready() first (download() does): the first call
// waits for them and frees their scratch.
struct DeviceDense {
  void* p = nullptr;
  i64 n = 0;
  DeviceDense() = default;
  DeviceDense(const DeviceDense&) = delete;
  DeviceDense& operator=(const DeviceDense&) = delete;
  ~DeviceDense();                        // waits for the kernels, frees the device memory
  void ready() const;                    // blocks until p holds the inverse (throws on a device error)
  std::vector<double> download() const;  // the n x n host copy
  // the computation still in flight: its stream (a hipStream_t) and scratch
  mutable void* pending = nullptr;
  mutable std::vector<void*> scratch;
  mutable std::mutex mu;
};

// Reduced-system factor (rounds of independent eliminations + dense core).
struct Factor {
  i64 n = 0;
  std::vector<u32> round_ptr, order;
  std::vector<double> pivot;
  std::vector<u32> col_ptr, col_idx;
  std::vector<double> col_w;
  std::vector<u32> core;
  std::vector<double> core_inv;  // |core|^2, anchored (zero last row / column)
  std::shared_ptr<DeviceDense> core_dev;  // or the same matrix in device memory (core_inv empty)
};

struct Subdomain {
  int q = 0, r = 0;
  std::vector<u32> gidx;    // q-levels then r-levels
  std::vector<double> pou;  // empty = uniform
  CrossTab ct;
  bool eliminate_q = true, dense = false;
  std::vector<double> inv_diag_elim;
  Factor f;                 // dense path: rounds empty, core = all kept vertices
  i64 n_local() const { return ct.n_q() + ct.n_r(); }
  i64 n_keep() const { return eliminate_q ? ct.n_r() : ct.n_q(); }
};

// Concatenated, device-ready preconditioner arrays.
struct PrecondFlat {
  i64 n_loc = 0, n_red = 0, n_sub = 0;
  // local space (per subdomain: q entries then r entries)
  std::vector<u32> loc_gidx;
  std::vector<double> loc_scale;      // sign (r-block negated) x partition weight
  std::vector<u32> loc_ptr;           // n_sub + 1
  std::vector<u32> red_ptr;           // n_sub + 1 (reduced = kept vertices)
  std::vector<u32> loc_sub;           // local entry -> subdomain
  std::vector<u32> red_sub;           // reduced entry -> subdomain
  std::vector<u32> red_loc;           // reduced entry -> its local entry
  // elimination rows: eliminated local entry e, inv diag, row of C / C' into reduced
  std::vector<u32> elim_loc;          // n_elim
  std::vector<double> elim_inv;       // n_elim
  std::vector<u32> elim_ptr, elim_idx;   // CSR: elim row -> reduced entries
  std::vector<double> elim_val;
  // keep rows: reduced entry j, CSR over eliminated local entries with value
  // C[e,j] * inv_diag[e], plus the row sum of those values
  std::vector<u32> keep_ptr, keep_idx;
  std::vector<double> keep_val, keep_rowsum;
  // factor levels in the reduced space (all subdomains merged by round)
  std::vector<u32> lvl_ptr;           // n_levels + 1 over lvl_rows
  std::vector<u32> lvl_rows;          // reduced entries eliminated per round
  std::vector<u32> fin_ptr, fin_idx;  // forward: in-edges (earlier vertex a, coef w/d_a)
  std::vector<double> fin_val;        //   indexed by position in lvl_rows
  std::vector<u32> bout_ptr, bout_idx;  // backward: out-edges (later vertex u, coef w/d_v)
  std::vector<double> bout_val;
  std::vector<double> inv_pivot;      // per reduced entry (0 for core / anchors)
  // dense cores
  std::vector<u32> core_ptr;          // n_cores + 1 over core_idx
  std::vector<u32> core_idx;          // reduced entries
  std::vector<uint64_t> core_mat_off; // n_cores
  std::vector<double> core_mat;       // concatenated row-major inverses
  std::vector<u32> core_row_core;     // per core row -> core id
  // per core: its inverse in device memory (null: in core_mat); a device
  // core takes no space in core_mat (its offset is that of the next core)
  std::vector<std::shared_ptr<DeviceDense>> core_dev;
  // DOF combine: dof -> local entries (subdomain order)
  std::vector<u32> dof_ptr, dof_loc;
  i64 max_core = 0, ac_nnz = 0, c_nnz = 0, dense_subs = 0;
};

// What the device setup and the scheduling metrics need of each subdomain
// (kept also when the preconditioner comes from serialized bytes and the
// subdomain objects are not rebuilt): its factor pair, eliminated count,
// dense-core size, block sizes and cross-tab nonzeros.
struct SubMeta { int32_t q = 0, r = 0; i64 n_elim = 0, n_core = 0, n_q = 0, n_r = 0, c_nnz = 0; };

struct DeviceLayout;
struct Plan {
  Design d;
  LocalConfig cfg;
  int precond = 1;            // 0 none, 1 additive, 2 multiplicative
  bool explicit_op = false;
  bool pre_given = false;     // pre / meta (and gram for precond 2) came from bytes
  std::vector<FactorPass> passes;  // one per factor
  std::vector<Subdomain> subs;     // empty when pre_given
  std::vector<SubMeta> meta;       // one per subdomain
  PrecondFlat pre;
  Csr gram;                   // explicit Gramian (when needed)
  bool has_gram = false;
  // set by a caller that lays the plan out for the device: build_plan then
  // lays out the operator (numbering, passes, Gramian) on a host thread of
  // its own while the preconditioner is built, into lay_op
  bool layout_early = false;
  std::shared_ptr<DeviceLayout> lay_op;
};

// Serialized preconditioner (this library's own format, host byte order):
// the preconditioner kind, local config, DOF count, per-subdomain metadata,
// the flattened arrays and, for multiplicative Schwarz, the Gramian its
// residual updates use.  Everything a device apply needs, in the user's DOF
// numbering.
std::vector<uint8_t> serialize_precond(const Plan& p);
// Fills p.precond, p.cfg, p.meta, p.pre (and p.gram), sets p.pre_given and
// returns the serialized DOF count (throws wa::Error on malformed bytes).
i64 deserialize_precond(const uint8_t* buf, i64 len, Plan& p);

// Builds everything the solver uploads (throws wa::Error).
void build_plan(Plan& p);

// The device works in an internal DOF numbering: only active levels (levels
// with at least one observation; the others are zero rows and columns of G
// and stay zero in every solve), factor blocks in factor order, the levels of
// the factor with the most levels ordered by the smallest code they meet of
// the factor with the next-most levels (workers grouped by firm), the rest in
// level order.  Every device array indexed by DOF is built in this numbering;
// the C ABI maps user DOF vectors in and out.  The local layout of each
// subdomain's eliminated block follows the same order.
struct DeviceLayout {
  i64 m = 0;                        // internal DOFs
  std::vector<u32> int_of;          // user DOF -> internal (UINT32_MAX = inactive)
  std::vector<u32> ext_of;          // internal -> user DOF
  std::vector<i64> off, levels;     // per factor: internal offset and active levels
  std::vector<u32> dof_of_obs;      // Q x n, factor-major: internal DOF of each observation
  std::vector<FactorPass> passes;   // rows = active levels in internal order, `other` internal
  Csr gram;                         // explicit Gramian in internal numbering (if built)
  PrecondFlat pre;                  // local / DOF arrays in internal numbering
};
DeviceLayout make_device_layout(const Plan& p);
// the same in two parts: the operator (numbering, passes, Gramian), then the
// preconditioner arrays (p.pre) on top of it
DeviceLayout layout_operator(const Plan& p);
void layout_precond(const Plan& p, DeviceLayout& L);

// Individual steps (exposed for tests through the C ABI).
std::vector<Subdomain> build_subdomains(const Design& d);
void build_local_solver(Subdomain& s, const LocalConfig& cfg);
Csr build_gramian(const Design& d, const std::vector<FactorPass>& passes);
std::vector<double> anchored_inverse(const std::vector<double>& a, i64 n);
// The same inverse evaluated on the GPU (dense.hip), bit for bit, and left in
// device memory, of the n x n Laplacian given as CSR (row i: neighbours col
// with positive weights w, diagonal diag_i): false when the anchored minor is
// not positive definite; throws wa::Error on a HIP failure.
bool device_anchored_inverse(i64 n, const u32* ptr, const u32* col, const double* w, const double* diag,
                             std::shared_ptr<DeviceDense>& inv);
// core c's inverse in device memory, or null
inline const DeviceDense* device_core(const PrecondFlat& f, size_t c) {
  return c < f.core_dev.size() ? f.core_dev[c].get() : nullptr;
}
// Copies device cores into core_mat (offsets rebuilt), so the arrays hold
// every core on the host (serialisation).
void materialize_cores(PrecondFlat& f);

}  // namespace wa
